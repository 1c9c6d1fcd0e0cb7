/*
 * wndgpu.h -- C ABI of the MI355X WND-CHARM feature engine.
 *
 * Plain pointers and sizes only.  Every function returns WNDGPU_OK (0) or a
 * negative error code; wndgpu_last_error() gives the message of the last
 * failure on that context.  One context per host thread; a context owns one
 * GPU stream, its device buffers and the host-computed basis tables.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   wndgpu_plan_standard        StdFeatureComputationPlans::getFeatureSet / getFeatureSetLong /
 *                               getFeatureSetColor / getFeatureSetLongColor
 *                               (src/Tasks.cpp:313-342, plans 354-508)
 *   wndgpu_plan_from_groups     FeatureComputationPlan(name) + add(group name) + finalize()
 *                               (src/Tasks.cpp:118-158, src/FeatureNames.cpp:309-396)
 *   wndgpu_plan_n_features / wndgpu_plan_feature_name
 *                               FeatureComputationPlan::n_features / getFeatureNameByIndex
 *                               (src/Tasks.h:166-170, src/Tasks.cpp:160-168)
 *   wndgpu_compute              FeatureComputationPlanExecutor::run(src, feature_mat, row)
 *                               (src/Tasks.cpp:275-295), for a whole batch of images
 *   wndgpu_transform / wndgpu_transform_image
 *                               ImageTransform::execute(in, out) (src/ImageTransforms.h:45-51,
 *                               src/ImageTransforms.cpp:63-109) by transform name
 *   wndgpu_algorithm            FeatureAlgorithm::execute(in) (src/FeatureAlgorithms.h:280-290,
 *                               src/FeatureAlgorithms.cpp:54-489) by algorithm name
 *   wndgpu_normalize_train      TrainingSet::normalize (src/wndchrm_src/TrainingSet.cpp:2090-2123)
 *   wndgpu_normalize_test       signatures::normalize (src/wndchrm_src/signatures.cpp:196-215)
 *   wndgpu_fisher               TrainingSet::SetFisherScores (TrainingSet.cpp:2183-2324)
 *   wndgpu_wnd5                 TrainingSet::classify2 (TrainingSet.cpp:2429-2586)
 *   wndgpu_wnn                  TrainingSet::WNNclassify (TrainingSet.cpp:2375-2417)
 *   wndgpu_py_normalize / wndgpu_py_fisher / wndgpu_py_wnd5
 *                               the wndcharm Python package's normalize_by_columns
 *                               (wndcharm/utils.py:124-240), FisherFeatureWeights.
 *                               NewFromFeatureSpace (FeatureWeights.py:165-236) and
 *                               NewWND5 (FeatureSpacePrediction.py:641-799)
 *   wndgpu_compute_sampled      TrainingSet::AddImageFile's sampling (TrainingSet.cpp:1425-1480):
 *                               OpenImage pre-processing, rotations, tiles
 *   wndgpu_tiff_read            ImageMatrix::LoadTIFF (src/cmatrix.cpp:85-203), the file part
 *   wndgpu_sig_write / _read    signatures::SaveToFile / LoadFromFilep
 *                               (src/wndchrm_src/signatures.cpp:227-304)
 *   wndgpu_fit_write / _read    TrainingSet::SaveToFile / ReadFromFile (TrainingSet.cpp:339-480)
 *   wndgpu_weights_write / _read
 *                               TrainingSet::SaveWeightVector / LoadWeightVector
 *                               (TrainingSet.cpp:594-642)
 */
#ifndef WNDGPU_H
#define WNDGPU_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum {
  WNDGPU_OK = 0,
  WNDGPU_E_ARG = -1,      /* bad argument (null pointer, size out of range, unknown name) */
  WNDGPU_E_DEVICE = -2,   /* HIP runtime error */
  WNDGPU_E_NOMEM = -3,    /* device allocation failed */
  WNDGPU_E_IMAGE = -4,    /* at least one image of the batch was rejected (see status[]) */
  WNDGPU_E_IO = -5,       /* a file could not be opened, read or written (wndgpu_io_error) */
  WNDGPU_E_FORMAT = -6    /* a file is not in a format this library reads (wndgpu_io_error) */
};

/* sample types.  RGB8 / RGB16: three interleaved samples per pixel (R, G, B),
 * `stride` counted in pixels; the image becomes the gray plane RGB2GRAY plus
 * the colour plane RGB2HSV (ImageMatrix::LoadTIFF with 3 samples per pixel,
 * src/cmatrix.cpp:134-193).  16-bit samples are scaled to bytes by the global
 * min / max over the three channels. */
enum { WNDGPU_U16 = 0, WNDGPU_F32 = 1, WNDGPU_F64 = 2, WNDGPU_U8 = 3, WNDGPU_RGB8 = 4, WNDGPU_RGB16 = 5 };

/* per-image status inside a batch */
enum {
  WNDGPU_IMG_OK = 0,
  WNDGPU_IMG_TOO_SMALL = 1,
  WNDGPU_IMG_TOO_LARGE = 2,
  WNDGPU_IMG_BAD_DTYPE = 3,
  WNDGPU_IMG_NOT_COLOR = 4, /* the plan reads the colour plane and the image is gray */
  WNDGPU_IMG_BAD_SAMPLING = 5 /* crop or tile outside the image, or normalisation of a float image */
};

/* smallest / largest side accepted (the Symlet-5 and Tamura paths need >= 16) */
#define WNDGPU_MIN_SIDE 16
#define WNDGPU_MAX_SIDE 4096

typedef struct wndgpu_ctx wndgpu_ctx;
typedef struct wndgpu_plan wndgpu_plan;

/* one input image: row-major, `stride` elements between rows */
typedef struct {
  const void *data;
  int height, width, stride;
  int dtype; /* WNDGPU_U16 / F32 / F64 / U8 / RGB8 / RGB16 */
} wndgpu_image;

int wndgpu_create(int device, wndgpu_ctx **ctx);
void wndgpu_destroy(wndgpu_ctx *ctx);
const char *wndgpu_last_error(const wndgpu_ctx *ctx);
const char *wndgpu_version(void);

/* plans: set 0 = short (1059 features), 1 = long (2919), 2 = short colour
 * (2199), 3 = long colour (4059); the colour sets need RGB images */
int wndgpu_plan_standard(int set, wndgpu_plan **plan);
/* custom plan from feature-group names such as "Haralick Textures (Fourier ())" */
int wndgpu_plan_from_groups(const char *const *group_names, int n_groups, wndgpu_plan **plan);
void wndgpu_plan_destroy(wndgpu_plan *plan);
int wndgpu_plan_n_features(const wndgpu_plan *plan);
const char *wndgpu_plan_feature_name(const wndgpu_plan *plan, int column);
/* Arithmetic of the basis contractions (the mixed-precision form; no
 * counterpart in the reference, which computes in FP64): WNDGPU_FP64
 * (default, the plain form), WNDGPU_FP32 or WNDGPU_BF16X3.  The reduced
 * forms run the Gabor filter bank ("Gabor Textures") in FP32 fused
 * multiply-adds over groups of 8 taps, or in bf16 split products on the
 * matrix cores (pixels split exactly into two bf16 parts, weights into
 * three, five products per tap, FP32 chunks of 80 products), the chunks
 * summed in FP64.  Each energy carries an a-priori error bound, and every
 * count that its interval leaves open (maximum, minimum, Otsu bin,
 * threshold, base count) is taken on the energy recomputed in FP64 in the
 * plain form's order: integer-valued features equal the plain form's.  The
 * Chebyshev and Chebyshev-Fourier histograms and the Zernike sums stay FP64
 * in every precision.  Set the precision on a plan of your own
 * (wndgpu_plan_from_groups): the standard plans are meant to stay plain. */
enum { WNDGPU_FP64 = 0, WNDGPU_FP32 = 1, WNDGPU_BF16X3 = 2 };
int wndgpu_plan_set_precision(wndgpu_plan *plan, int precision);
/* column offset of a feature group in the plan, or -1 */
int wndgpu_plan_group_offset(const wndgpu_plan *plan, const char *group_name);

/* Features of a batch of host images.  out: host [n_images][n_features]
 * (row i written for every image whose status is WNDGPU_IMG_OK).  status may
 * be NULL.  Returns WNDGPU_E_IMAGE if any image was rejected. */
int wndgpu_compute(wndgpu_ctx *ctx, const wndgpu_plan *plan, const wndgpu_image *images, int n_images,
                   double *out, int *status);
/* Same, with image data and out already in device memory of the context's
 * GPU (pointers from hipMalloc / torch tensors); no host round trip. */
int wndgpu_compute_device(wndgpu_ctx *ctx, const wndgpu_plan *plan, const wndgpu_image *images, int n_images,
                          double *out_device, int *status);

/* Sampling of each image before the features, as wndchrm does it:
 * ImageMatrix::OpenImage's crop, Downsample and normalisation
 * (src/cmatrix.cpp:252-294), then rotations x tiles_x x tiles_y samples in
 * setup_featureset's order (src/wndchrm_src/wndchrm.cpp:89-128: rotation
 * outer, tile x, tile y inner), each the tile of the image rotated by
 * 90 * r degrees (Rotate, cmatrix.cpp:727-767, whose "270" is a transpose;
 * tile sizes as TrainingSet::AddImageFile, TrainingSet.cpp:1455-1480).
 * The colour plane follows the pixels (the reference does not rotate it).
 * A crop or tile that falls outside its image rejects the image
 * (WNDGPU_IMG_BAD_SAMPLING), where the reference's submatrix leaves the
 * previous tile in place. */
typedef struct {
  int roi_x, roi_y, roi_w, roi_h; /* bounding rectangle; roi_x < 0: the whole image */
  int downsample;                 /* percent, applied when 0 < downsample < 100 */
  double mean, stdev;             /* normalise to mean (when > 0) and stdev (when > 0): 8/16-bit images */
  int rotations;                  /* 1..4 */
  int tiles_x, tiles_y;           /* >= 1 */
} wndgpu_sampling;
int wndgpu_sampling_n_samples(const wndgpu_sampling *sampling);
/* out: host [n_images * n_samples][n_features], sample k of image i at row
 * i * n_samples + k; status per image.  Images may be up to 32768 a side;
 * each sample must be within [WNDGPU_MIN_SIDE, WNDGPU_MAX_SIDE]. */
int wndgpu_compute_sampled(wndgpu_ctx *ctx, const wndgpu_plan *plan, const wndgpu_image *images, int n_images,
                           const wndgpu_sampling *sampling, double *out, int *status);

/* One transform ("Fourier", "Chebyshev", "Wavelet", "Edge", "Color
 * Transform") on one FP64 host plane.  out must hold the output plane: Fourier/Edge h*w, Chebyshev
 * n*n with n = min(h,w), Wavelet 2*floor((h+9)/2) * 2*floor((w+9)/2). */
int wndgpu_transform(wndgpu_ctx *ctx, const char *name, const double *in, int h, int w, double *out, int *out_h,
                     int *out_w);
/* One transform of one host image of any sample type, by name: the gray
 * transforms above, "Color Transform" (the identity copy of the gray plane),
 * and "Color" / "Hue", which read the colour plane of an RGB image
 * (ColorTransform / HueTransform::execute, src/ImageTransforms.cpp:119-136,
 * src/cmatrix.cpp:1023-1066).  out as for wndgpu_transform. */
int wndgpu_transform_image(wndgpu_ctx *ctx, const char *name, const wndgpu_image *image, double *out, int *out_h,
                           int *out_w);
/* One feature algorithm by reference name ("Zernike Coefficients", ...) on
 * one FP64 host plane; out gets n_features doubles (returned in *n_out). */
int wndgpu_algorithm(wndgpu_ctx *ctx, const char *name, const double *in, int h, int w, double *out, int *n_out);
/* n_features of an algorithm name, or -1 */
int wndgpu_algorithm_n_features(const char *name);

/* Classifier (host arrays, FP64, row-major [n_samples][n_features];
 * class labels 1..n_classes, 0 for a sample of no class; a label outside
 * 0..n_classes is WNDGPU_E_ARG). */
int wndgpu_normalize_train(wndgpu_ctx *ctx, double *data, int n_samples, int n_features, double *mins,
                           double *maxs);
int wndgpu_normalize_test(wndgpu_ctx *ctx, double *vec, int n_features, const double *mins, const double *maxs);
int wndgpu_fisher(wndgpu_ctx *ctx, const double *data, const int *classes, int n_samples, int n_features,
                  int n_classes, double used_fraction, double *weights);
/* WND5 marginal probabilities for n_test already-normalised test vectors:
 * probabilities [n_test][n_classes], norm_factor [n_test], predicted [n_test]. */
int wndgpu_wnd5(wndgpu_ctx *ctx, const double *train, const int *classes, int n_samples, int n_features,
                int n_classes, const double *weights, const double *test, int n_test, double *probabilities,
                double *norm_factor, long *predicted);
/* Weighted nearest neighbour (WNNclassify) for n_test already-normalised
 * test vectors, same arrays: distance sqrt(sum_s w_s (x_s - y_s)^2)
 * (TrainingSet::distance with power 2), an identical training vector
 * (distance below 1/INF) skipped, the class of the first closest sample
 * predicted (0 if none), probabilities[c] = (1/d_c) / sum_c' (1/d_c') with
 * d_c the closest distance of class c, norm_factor = that sum.  The
 * reference's skip of a training sample with the test sample's file path
 * has no counterpart (vectors carry no paths). */
int wndgpu_wnn(wndgpu_ctx *ctx, const double *train, const int *classes, int n_samples, int n_features,
               int n_classes, const double *weights, const double *test, int n_test, double *probabilities,
               double *norm_factor, long *predicted);

/* The same classifier as the wndcharm Python package computes it (its maths
 * differs from TrainingSet's; host arrays, FP64, row-major [n][F]):
 *   wndgpu_py_normalize  normalize_by_columns (wndcharm/utils.py:172-240),
 *                        with ReplaceNonReal (utils.py:124-169): flags
 *                        WNDGPU_PY_ZSCORE (else min-max to 0..100),
 *                        WNDGPU_PY_NON_REAL_CHECK, WNDGPU_PY_GIVEN_* (those
 *                        arrays are reference values, read; the others of
 *                        the chosen form are written), WNDGPU_PY_REPLACE_ONLY
 *                        (ReplaceNonReal alone).  *replaced (may be NULL):
 *                        whether the matrix held a NaN or +/-inf.
 *   wndgpu_py_fisher     FisherFeatureWeights.NewFromFeatureSpace
 *                        (wndcharm/FeatureWeights.py:165-236): ddof=1 class
 *                        variances, no floor, NaN / inf weights -> 0; rows of
 *                        class 0 count in the population mean only.
 *   wndgpu_py_wnd5       FeatureSpaceClassification.NewWND5
 *                        (wndcharm/FeatureSpacePrediction.py:641-799): weighted
 *                        features, squared euclidean distances, distances
 *                        <= DBL_EPSILON masked, per-class mean of d^-5; a class
 *                        with no unmasked sample gives no prediction (NaN
 *                        probabilities and norm factor, predicted 0). */
enum { WNDGPU_PY_ZSCORE = 1, WNDGPU_PY_NON_REAL_CHECK = 2, WNDGPU_PY_GIVEN_MINS = 4, WNDGPU_PY_GIVEN_MAXS = 8,
       WNDGPU_PY_GIVEN_MEANS = 16, WNDGPU_PY_GIVEN_STDEVS = 32, WNDGPU_PY_REPLACE_ONLY = 64 };
int wndgpu_py_normalize(wndgpu_ctx *ctx, double *data, int n_samples, int n_features, int flags, double *mins,
                        double *maxs, double *means, double *stdevs, int *replaced);
int wndgpu_py_fisher(wndgpu_ctx *ctx, const double *data, const int *classes, int n_samples, int n_features,
                     int n_classes, double *weights);
int wndgpu_py_wnd5(wndgpu_ctx *ctx, const double *train, const int *classes, int n_samples, int n_features,
                   int n_classes, const double *weights, const double *test, int n_test, double *probabilities,
                   double *norm_factor, long *predicted);

/* ---- Differentiable form: forward and adjoint (reverse-mode) passes of the
 * smooth transforms and features, and of the WND similarity, for gradients
 * of class probabilities with respect to pixels.  No counterpart in the
 * reference, which has no derivatives; the forward values follow:
 *   WNDGPU_AD_FOURIER          FourierTransform::execute -> ImageMatrix::fft2
 *                              (src/ImageTransforms.cpp:161-165, src/cmatrix.cpp:1107-1200): h x w -> h x w
 *   WNDGPU_AD_CHEBYSHEV        ChebyshevTransform::execute (src/ImageTransforms.cpp:175-179,
 *                              src/cmatrix.cpp:1203-1222), order min(h,w): h x w -> N x N
 *   WNDGPU_AD_CHEBYSHEV_COEFFS the order-20 coefficients that ChebyshevCoefficients
 *                              histograms (src/FeatureAlgorithms.cpp:73-93, cmatrix.cpp:1274-1281): -> 20 x 20
 *   WNDGPU_AD_ZERNIKE          ZernikeCoefficients::execute (src/FeatureAlgorithms.cpp:98-116,
 *                              src/textures/zernike/zernike.cpp:207-341): -> 72 values
 *   WNDGPU_AD_CHEB_FOURIER     the NN x NN coefficient magnitudes that ChebyshevFourierCoefficients
 *                              histograms, after its Downsample above 300 x 300 pixels
 *                              (src/FeatureAlgorithms.cpp:54-69, src/cmatrix.cpp:579-720,1227-1240,
 *                              src/transforms/ChebyshevFourier.cpp:59-160): -> NN*NN values (NN <= 23)
 *   WNDGPU_AD_COMB             the line moments that CombFirstFourMoments histograms
 *                              (src/FeatureAlgorithms.cpp:184-200, src/statistics/CombFirst4Moments.cpp:78-217):
 *                              -> 320 values, [comb][mean, std, skew, kurt][20 positions]
 *   WNDGPU_AD_GABOR            the 8 energy maps GaborTextures thresholds (GaborEnergy,
 *                              src/textures/gabor.cpp:166-231; f0 = 0.1, then 1..7): -> (8 h) x w
 * All pointers are device memory of the context's GPU: in [n][h][w] FP64
 * planes of one size, out / grad_out [n][out_h][out_w], grad_in [n][h][w]
 * (overwritten).  work: caller-owned device scratch of
 * wndgpu_ad_work_bytes(op, n, h, w) bytes.  stream: the hipStream_t to issue
 * on (NULL: the default stream), ordered after the caller's earlier work on
 * it; calls return without waiting.  The backward pass recomputes what it
 * needs from `in`. */
enum { WNDGPU_AD_FOURIER = 0, WNDGPU_AD_CHEBYSHEV = 1, WNDGPU_AD_ZERNIKE = 2, WNDGPU_AD_CHEBYSHEV_COEFFS = 3,
       WNDGPU_AD_CHEB_FOURIER = 4, WNDGPU_AD_COMB = 5, WNDGPU_AD_GABOR = 6 };
int wndgpu_ad_output_shape(int op, int h, int w, int *out_h, int *out_w);
size_t wndgpu_ad_work_bytes(int op, int n, int h, int w);
int wndgpu_ad_forward(wndgpu_ctx *ctx, int op, const double *in, int n, int h, int w, double *out, void *work,
                      void *stream);
int wndgpu_ad_backward(wndgpu_ctx *ctx, int op, const double *in, const double *grad_out, int n, int h, int w,
                       double *grad_in, void *work, void *stream);
/* WND5 marginal probabilities of n_test raw feature vectors x [n_test][F]
 * (normalised inside with the training mins / maxs, as
 * signatures::normalize, src/wndchrm_src/signatures.cpp:196-215, then
 * TrainingSet::classify2, TrainingSet.cpp:2429-2586) and their adjoint:
 * grad_x = (dprob/dx)^T grad_prob.  train [n_samples][F] normalised,
 * classes 1..n_classes (int32; the kernels skip a sample whose label lies
 * outside 0..n_classes, which the caller is to prevent), weights / mins /
 * maxs [F]; all device memory; work of wndgpu_ad_wnd_work_bytes bytes. */
size_t wndgpu_ad_wnd_work_bytes(int n_samples, int n_features, int n_classes, int n_test);
int wndgpu_ad_wnd_forward(wndgpu_ctx *ctx, const double *train, const int *classes, int n_samples, int n_features,
                          int n_classes, const double *weights, const double *mins, const double *maxs,
                          const double *x, int n_test, double *prob, void *work, void *stream);
int wndgpu_ad_wnd_backward(wndgpu_ctx *ctx, const double *train, const int *classes, int n_samples, int n_features,
                           int n_classes, const double *weights, const double *mins, const double *maxs,
                           const double *x, int n_test, const double *grad_prob, double *grad_x, void *work,
                           void *stream);

/* wndgpu_ad_wnd_forward without the work argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * work lives in a file-scope device array of WNDGPU_NOALLOC_SCRATCH_BYTES
 * that each GPU holds from the library's load until the process ends; its
 * address is looked up once per device by wndgpu_create.  The call needs
 * 8 * n_test * (n_features + n_samples + n_classes + 2) bytes of it and
 * returns WNDGPU_E_ARG above the maximum (split the test vectors).  It calls
 * no allocating routine.  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (other streams, other host threads, other
 * contexts) race, so order them.  Inputs, output and stream as for
 * wndgpu_ad_wnd_forward. */
#define WNDGPU_NOALLOC_SCRATCH_BYTES (32u << 20)
int wndgpu_ad_wnd_forward_noalloc(wndgpu_ctx *ctx, const double *train, const int *classes, int n_samples,
                                  int n_features, int n_classes, const double *weights, const double *mins,
                                  const double *maxs, const double *x, int n_test, double *prob, void *stream);

/* Batch graphs (no counterpart in the reference, whose executor runs one
 * image at a time).  mode 0 (default): every batch issues its launches.
 * mode 1: a batch whose signature (plan, image geometry, sampling, and the
 * context's arena and table state) was seen before is captured into a
 * hipGraph once, then replayed: one graph launch per call, with the new
 * image and output addresses flowing through the staged descriptors that
 * the graph's first node copies.  mode 2: as 1, and the two launch
 * schedules (the Haralick chain on a side stream, or every launch on one
 * stream) are both captured, timed on their first replays and the faster
 * kept.  Graphs are dropped when the context's arena grows or a table is
 * freed.  WNDGPU_GRAPH=1 / 2 sets the mode at context creation.
 * wndgpu_graph_stats: captures and replays so far, graphs cached, and the
 * schedule picked for the most recently used graph (-1 while undecided). */
int wndgpu_set_graph_mode(wndgpu_ctx *ctx, int mode);
int wndgpu_graph_stats(const wndgpu_ctx *ctx, long *captures, long *replays, int *cached, int *schedule_of_last);

/* Per-launch timings of the last wndgpu_compute / wndgpu_compute_device made
 * with WNDGPU_TIMING set in the environment (ms from a HIP event pair around
 * each launch, in execution order; launches are named by feature group, and
 * the Haralick group by step), the counterpart of the reference's TIME lines
 * (src/Tasks.cpp:249-253).  Returns the number of entries written. */
int wndgpu_last_node_times(wndgpu_ctx *ctx, double *ms, const char **names, int cap);

/* Work record of the Haralick groups of the last compute, for roofline
 * accounting: tones[k] = Ng of Haralick entry k (one (image, node) pair),
 * schur_flops[4k+a] = floating-point operations the Francis QR issued for
 * angle a.  Returns the number of entries (<= cap are written).  No
 * counterpart in the reference (Ng is local to CVIPtexture.cpp:400-414). */
int wndgpu_last_haralick_work(wndgpu_ctx *ctx, int *tones, double *schur_flops, int cap);

/* Chebyshev-Fourier entries of the last compute (one (image, node) pair
 * each, in the order of the plan's entries for that group): open[k] = 1
 * where the histogram of the matrix-core sums could not be certified equal
 * to the reference's and the entry was recomputed in the reference's order
 * (ChebyshevFourier.cpp:112-135).  Returns the number of entries (<= cap
 * are written).  WNDGPU_CF_SLACK (a factor >= 1, read at context creation)
 * widens the certification bound.  No counterpart in the reference. */
int wndgpu_last_cf_open(wndgpu_ctx *ctx, int *open, int cap);

/* Size-dependent tables ahead of time: runs the sizing pass of a batch that
 * a later wndgpu_compute / wndgpu_compute_device will be given and queues the
 * Chebyshev bases and Chebyshev-Fourier geometry tables it lacks (glibc libm,
 * as ChebyshevFourier.cpp:59-103 and chebyshev.cpp compute them) on a pool
 * of up to 16 host threads, then returns.  The compute call that needs a
 * queued table waits for that table only; tables no compute takes within two
 * calls are dropped.  Pixels are not read (images only need their sizes and
 * types).  Meant for a loop that prefetches batch k+1 before computing batch
 * k, so the host work overlaps the device's.  No counterpart in the reference
 * (its tables are per-call locals). */
int wndgpu_prefetch_tables(wndgpu_ctx *ctx, const wndgpu_plan *plan, const wndgpu_image *images, int n_images);

/* Wall times (ms) of the last compute call: host_ms from entry until every
 * launch was issued (sizing, tables, descriptors), table_ms the part of it
 * spent building or awaiting size-dependent tables, device_ms waiting for the
 * device afterwards. */
int wndgpu_last_call_times(const wndgpu_ctx *ctx, double *host_ms, double *table_ms, double *device_ms);

/* ---- files (host only: no context, no GPU).  Errors: WNDGPU_E_IO /
 * WNDGPU_E_FORMAT / WNDGPU_E_ARG with a message in wndgpu_io_error() of the
 * calling thread. */
const char *wndgpu_io_error(void);
void wndgpu_free(void *p);

/* A TIFF image as LoadTIFF reads it: one directory, strips (uncompressed,
 * LZW, PackBits or Deflate, optional horizontal predictor), 8 or 16 bits per
 * sample, 1 sample per pixel (dtype WNDGPU_U8 / U16) or 3 (WNDGPU_RGB8 /
 * RGB16).  *pixels is allocated here (free with wndgpu_free); image->data
 * points into it, ready for wndgpu_compute. */
int wndgpu_tiff_read(const char *path, void **pixels, wndgpu_image *image);

/* .sig: "class<TAB>version.type" (or "%f" of sample_value when it is not
 * NULL), the image path, then "%.8g<TAB>name" per feature ("%.8g" when
 * names is NULL). */
int wndgpu_sig_write(const char *path, const char *image_path, int sample_class, const double *sample_value,
                     int version, int vec_type, const double *values, const char *const *names, int n);

/* Feature tables read from files: one sample (.sig) or a training set (.fit) */
typedef struct wndgpu_table wndgpu_table;
/* continuous: read the first field as a value (class 1), as for a continuous training set */
int wndgpu_sig_read(const char *path, int continuous, wndgpu_table **table);
int wndgpu_fit_read(const char *path, wndgpu_table **table);
int wndgpu_table_shape(const wndgpu_table *table, int *n_samples, int *n_features, int *n_classes, int *version,
                       int *vec_type);
const double *wndgpu_table_values(const wndgpu_table *table);        /* [n_samples][n_features] */
const int *wndgpu_table_classes(const wndgpu_table *table);          /* [n_samples] */
const double *wndgpu_table_sample_values(const wndgpu_table *table); /* [n_samples] */
const char *wndgpu_table_feature_name(const wndgpu_table *table, int feature);
const char *wndgpu_table_class_label(const wndgpu_table *table, int cls); /* 0..n_classes */
const char *wndgpu_table_sample_path(const wndgpu_table *table, int sample);
void wndgpu_table_destroy(wndgpu_table *table);

/* .fit: header, feature names, n_classes + 1 class labels (label 0 is the
 * unknown class), then per sample its values ("%ld" when integral, else
 * "%.8g", space-terminated), its class ("%f" of sample_values when that is
 * not NULL) and its path (empty when sample_paths is NULL). */
int wndgpu_fit_write(const char *path, int n_classes, int version, int vec_type, int n_samples, int n_features,
                     const double *values, const int *classes, const double *sample_values,
                     const char *const *feature_names, const char *const *class_labels,
                     const char *const *sample_paths);

/* weights: "%f name" per feature.  The reader returns the number of weights
 * in *n_read (an error if above cap); negative weights read as 0. */
int wndgpu_weights_write(const char *path, const double *weights, const char *const *names, int n);
int wndgpu_weights_read(const char *path, double *weights, int cap, int *n_read);

#ifdef __cplusplus
}
#endif
#endif
