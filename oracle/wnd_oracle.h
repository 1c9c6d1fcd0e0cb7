/*
 * wnd_oracle.h -- CPU FP64 restatement of the WND-CHARM feature-extraction
 * path and the WND5 classifier.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * MI355X engine and the timed CPU baseline ("port").  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product library never links it and has no CPU fallback through it.
 *
 * Every entry point follows the reference's own arithmetic (same expression
 * order, FP64, built with -ffp-contract=off); the file:line of the reference
 * function each one restates is given in wnd_oracle.cpp.
 */
#ifndef WND_ORACLE_H
#define WND_ORACLE_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

/* transform ids (reference names: "Fourier", "Chebyshev", "Wavelet", "Edge",
 * "Color", "Hue"; "Color Transform" is the unregistered name of the standard
 * colour plans, an identity copy) */
enum { ORC_T_FOURIER = 0, ORC_T_CHEBYSHEV = 1, ORC_T_WAVELET = 2, ORC_T_EDGE = 3,
       ORC_T_COLOR = 4, ORC_T_HUE = 5, ORC_T_COLOR_TRANSFORM = 6 };

/* feature algorithm ids, with n_features in brackets */
enum {
  ORC_A_CHEB_FOURIER = 0,  /* Chebyshev-Fourier Coefficients [32] */
  ORC_A_CHEBYSHEV = 1,     /* Chebyshev Coefficients [32] */
  ORC_A_ZERNIKE = 2,       /* Zernike Coefficients [72] */
  ORC_A_HARALICK = 3,      /* Haralick Textures [28] */
  ORC_A_MSH = 4,           /* Multiscale Histograms [24] */
  ORC_A_TAMURA = 5,        /* Tamura Textures [6] */
  ORC_A_COMB = 6,          /* Comb Moments [48] */
  ORC_A_RADON = 7,         /* Radon Coefficients [12] */
  ORC_A_FRACTAL = 8,       /* Fractal Features [20] */
  ORC_A_PIXSTATS = 9,      /* Pixel Intensity Statistics [5] */
  ORC_A_EDGE = 10,         /* Edge Features [28] */
  ORC_A_OTSU = 11,         /* Otsu Object Features [34] */
  ORC_A_INV_OTSU = 12,     /* Inverse-Otsu Object Features [34] */
  ORC_A_GABOR = 13,        /* Gabor Textures [7] */
  ORC_A_GINI = 14,         /* Gini Coefficient [1] */
  ORC_A_COLOR_HIST = 15,   /* Color Histogram [20] (colour images) */
  ORC_A_COUNT = 16
};

int orc_alg_n_features(int alg);

/* Image transform: in is row-major H x W FP64.  out must hold the output
 * (Fourier/Edge: H*W; Chebyshev: N*N, N=min(W,H); Wavelet:
 * 2*floor((H+9)/2) * 2*floor((W+9)/2)).  Returns 0 on success. */
int orc_transform(int kind, const double *in, int h, int w, double *out, int *out_h, int *out_w);

/* Feature algorithm on one image; out gets orc_alg_n_features(alg) doubles. */
int orc_algorithm(int alg, const double *in, int h, int w, double *out);

/* Full plan.  set: 0 = short (1059), 1 = long (2919), 2 = short colour
 * (2199), 3 = long colour (4059).  out gets n features.  orc_features takes
 * a gray FP64 plane (sets 0, 1); orc_features_rgb an interleaved RGB image
 * with 8 or 16 bits per sample (any set). */
int orc_plan_n_features(int set);
int orc_features(int set, const double *img, int h, int w, double *out);
int orc_features_rgb(int set, const void *rgb, int bits, int h, int w, double *out);

/* Colour plane of an RGB image: gray = RGB2GRAY, hsv = RGB2HSV bytes. */
int orc_rgb_planes(const void *rgb, int bits, int h, int w, double *gray, unsigned char *hsv);
int orc_rgb_to_hsv(int r, int g, int b, unsigned char *hsv3);
long orc_find_color(int h, int s, int v); /* fuzzy colour class 1..19 */
void orc_find_color_many(const unsigned char *hsv, long n, long *out); /* n (h, s, v) byte triples */
/* kind ORC_T_COLOR or ORC_T_HUE of an RGB image; out gets H*W */
int orc_color_transform(int kind, const void *rgb, int bits, int h, int w, double *out);
int orc_color_histogram(const void *rgb, int bits, int h, int w, double *out20);
/* column -> feature-name string of a plan ("Zernike Coefficients (Fourier ()) [3]") */
const char *orc_feature_name(int set, int column);

/* Sampling of an image before the features: OpenImage's crop (roi_x < 0:
 * none), downsample (percent, when 0 < downsample < 100) and mean / stdev
 * normalisation (when mean > 0), then rotations (1..4) x tiles_x x tiles_y
 * samples, rotation outer, tile y inner.  dtype: 0 uint16, 2 float64,
 * 3 uint8, 4 RGB8, 5 RGB16 (row-major, RGB interleaved). */
typedef struct {
  int roi_x, roi_y, roi_w, roi_h, downsample;
  double mean, stdev;
  int rotations, tiles_x, tiles_y;
} orc_sampling;
int orc_sample_count(const orc_sampling *s);
int orc_sample_plane(const void *img, int dtype, int h, int w, const orc_sampling *s, int k, double *plane,
                     unsigned char *hsv, int *out_h, int *out_w);
/* out: [n_samples][n_features of set] */
int orc_features_sampled(int set, const void *img, int dtype, int h, int w, const orc_sampling *s, double *out);

/* Intermediates used by integer-exact parity tests. */
/* the NN x NN magnitudes ChebyshevFourierCoefficients histograms (after its
 * Downsample above 300 x 300); out: 529 doubles, *nn = NN */
int orc_cf_magnitudes(const double *in, int h, int w, double *out, int *nn);
/* the line moments CombFirstFourMoments histograms: [comb][mean, std, skew, kurt][20 positions] */
int orc_comb_raw(const double *in, int h, int w, double *out320);
/* the 8 Gabor energy maps of GaborTextureFilters2D (f0 = 0.1, 1..7): out [8][h][w] */
int orc_gabor_energy(const double *in, int h, int w, double *out);
double orc_otsu(const double *in, int h, int w);                         /* Otsu(dynamic) */
int orc_radon_proj(const double *in, int h, int w, double *out4xr, int *rsize);
int orc_haralick_grays(const double *in, int h, int w, unsigned char *out);
int orc_cooc(const unsigned char *grays, int h, int w, int angle, double *counts256x256, int *count);
unsigned long orc_bwlabel(const double *in, int h, int w, double *labels); /* Otsu mask + 8-conn labels */
int orc_haralick_f14(const double *P, int ng, double *eig_re, double *eig_im);

/* Classifier (reference TrainingSet.cpp / signatures.cpp). */
void orc_normalize_train(double *data, int n, int f, double *mins, double *maxs);
void orc_normalize_test(double *vec, int f, const double *mins, const double *maxs);
void orc_fisher(const double *data, const int *cls, int n, int f, int n_classes,
                double used, double *weights);
long orc_wnd5(const double *train, const int *cls, int n, int f, int n_classes,
              const double *weights, const double *test_normalized,
              double *probabilities, double *norm_factor);
long orc_wnn(const double *train, const int *cls, int n, int f, int n_classes,
             const double *weights, const double *test_normalized,
             double *probabilities, double *norm_factor);

/* Classifier as the wndcharm Python package computes it
 * (wndcharm/utils.py, FeatureWeights.py, FeatureSpacePrediction.py). */
double orc_np_sum(const double *a, long n); /* numpy's 1-D float64 sum */
int orc_py_normalize(double *data, int n, int f, int flags, double *mins, double *maxs, double *means,
                     double *stdevs);
void orc_py_fisher(const double *data, const int *cls, int n, int f, int n_classes, double *weights);
long orc_py_wnd5(const double *train, const int *cls, int n, int f, int n_classes, const double *weights,
                 const double *test, double *probabilities, double *norm_factor);

#ifdef __cplusplus
}
#endif
#endif
