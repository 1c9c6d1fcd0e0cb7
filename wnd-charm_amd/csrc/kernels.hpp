// Kernel descriptors and launcher declarations (host + device visible).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "common.hpp"

namespace wnd {

struct LoadDesc {
  const void *src;
  int h, w, stride;
  int dtype;             // 0 uint16, 1 float32, 2 float64, 3 uint8, 4 RGB 8-bit, 5 RGB 16-bit (interleaved)
  unsigned char *hsv;    // colour plane (H, S, V bytes per pixel) of an RGB image, else null
  const double *range;   // RGB 16-bit: min, max over the three channels
};

// C[r][c] = sum_a (A[r*sar + a*sak] * B[a*N + c]) / 2   (M x K) (K x N)
struct GemmDesc {
  const double *A;
  const double *B;
  double *C;
  int M, K, N;
  long sar, sak;
};

// C[b][r][c] = alpha * sum_a A[b*sab + r*sar + a*sak] * B[b*sbb + a*sbk + c*sbc],
// C row-major M x N at C + b*scb (adjoint path, k_adjoint.hip)
struct AdGemm {
  const double *A;
  long sab, sar, sak;
  const double *B;
  long sbb, sbk, sbc;
  double *C;
  long scb;
  int M, K, N;
  double alpha;
};

struct FftPlan {
  const double2 *tw;  // tw[k] = exp(-2 pi i k / n)
  int n, nst;
  int radix[20];
};

// Chebyshev-Fourier coefficient tables for one (possibly downsampled) size
struct CfDesc {
  const double *img;    // plane the coefficients are taken from
  const int *kk;        // in-disk pixel indices, column-major order (x outer)
  const double *tn;     // [np][NN] Chebyshev T(2r-1), T0 = 0.5
  const double *fc;     // [np][NN] Ftrm*cos(mf*phi)
  const double *fs;     // [np][NN] Ftrm*sin(-mf*phi)
  int np, NN;
  int w, h;             // plane size (row-major [h][w])
  double *sum;          // [2][NN*NN] real and imaginary coefficient sums (reference order)
  double *part;         // [nsplit][CFM_PART] matrix-core partial sums
  int nsplit;
};

// up to CF_GROUP entries whose planes share one set of basis tables (the
// reference-order recomputation of entries whose histogram stays open)
constexpr int CF_GROUP = 4;
struct CfGroup {
  const int *kk;
  const double *tn, *fc, *fs;
  int np, NN, ne;
  const double *img[CF_GROUP];
  double *sum[CF_GROUP];
  int ent[CF_GROUP];  // entry (output row) of each member
};

// up to CFM_GROUP entries of one table set for the matrix-core contraction;
// a launch splits each group's pixels into nsplit contiguous chunk ranges
constexpr int CFM_GROUP = 16;
constexpr int CFM_PART = 6 * 256 + 8;  // 3x2 16x16 FP64 tiles as MFMA registers, then sum |I|
struct CfMGroup {
  const int *kk;
  const double *tn, *fc, *fs;
  int np, NN, ne;
  const double *img[CFM_GROUP];
  double *part[CFM_GROUP];
};

struct HsvDsDesc {  // Downsample of a colour plane (3 bytes per pixel), positions as DsDesc
  const unsigned char *in;
  unsigned char *tmp;  // [oh][ow] after the x pass
  unsigned char *out;  // [nh][nw]
  int oh, ow, nh, nw, cx, cy;
  const double *xs;
  const double *ys;
  double dx, dy;
};

struct NormDesc {  // normalize to a mean / stdev (cmatrix.cpp:836-888), in place
  double *p;
  long n;
  double mean, stdev, max_range;
  double d_mean, std_fact;  // written by the statistics pass
};

struct SampleDesc {  // one rotated, tiled sample of a pre-processed plane
  const double *src;
  const unsigned char *hsrc;  // colour plane of the source, or null
  unsigned char *hdst;        // colour plane of the sample, or null
  int sh, sw, rot, x0, y0;
};

struct DsDesc {  // area-weighted Downsample (cmatrix.cpp:579-720)
  const double *in;
  double *tmp;          // [oh][ow] after the x pass
  double *out;          // [nh][nw]
  int oh, ow, nh, nw, cx, cy;  // cx/cy: number of x / y steps
  const double *xs;     // x positions (cx+1 entries, sequential accumulation)
  const double *ys;
  double dx, dy;
};

void launch_load(const LoadDesc *d, Img *dst, int nimg, long maxn, hipStream_t s);
void launch_stats(const Img *imgs, Stats *st, int nimg, hipStream_t s);
void launch_color_lut(unsigned char *lut, hipStream_t s);
void launch_color_node(const unsigned char *const *hsv, const unsigned char *lut, const Img *out, int kind, int nimg,
                       long maxp, hipStream_t s);
void launch_color_hist(const unsigned char *const *hsv, const Img *imgs, const unsigned char *lut, OutRef o, int nimg,
                       hipStream_t s);
void launch_rgb_range(const LoadDesc *d, double *range, int nimg, hipStream_t s);
void launch_gather_stats(const Stats *const *src, Stats *dst, int n, hipStream_t s);
void launch_edge(const Img *in, const Img *out, int nimg, long maxn, hipStream_t s);
void launch_wavelet(const Img *in, double *const *g, const int *gstride, const Img *out, int nimg, int maxh,
                    int maxw, hipStream_t s);
void launch_cheb_gemm(const GemmDesc *d, int nimg, int maxM, int maxN, hipStream_t s);
void launch_fft(const Img *in, double2 *const *f1, const FftPlan *plw, const FftPlan *plh, const Img *out, int nimg,
                int maxh, int maxw, hipStream_t s);
void launch_sort(const Img *imgs, uint64_t *const *buf, const long *cap, int nimg, hipStream_t s);
void launch_median_gini(const Img *imgs, uint64_t *const *buf, const Stats *st, double *median_out, OutRef pix,
                        OutRef gini, int nimg, hipStream_t s);
void launch_msh(const Img *imgs, const Stats *st, OutRef o, int nimg, hipStream_t s);
void launch_fractal(const Img *imgs, OutRef o, int nimg, int maxh, int maxw, hipStream_t s);
void launch_hist_values(const Img *vals, int nbins, OutRef o, int nimg, hipStream_t s);

// textures (k_texture.hip)
struct HarDesc {
  const Img *img;
  uint8_t *gray;        // [n] quantised plane
  uint32_t *cooc;       // [4][256*257/2] upper-triangle counts per angle
  double *P;            // [4][ng*ng] normalised matrices (ng <= 256)
  double *Q;            // [4][ng*ng]
  double *feat;         // [4][14]
  double *marg;         // [4][2][256] row / column marginals px, py
  double *work;         // [4] floating-point operations of the Schur iteration (reported, not used)
  int *ng;              // number of tones present
  int *lut;             // [256] tone -> rank
  int *count;           // [4] pair counts
};
// Haralick in four stream-ordered steps: quantise / co-occurrence / f1..f13
// / Q, Hessenberg reduction of Q, Francis QR (f14), mean and range output
void launch_haralick_prep(const Img *imgs, const Stats *st, const HarDesc *hd, int nimg, int maxn, hipStream_t s);
// order: the entries by decreasing ng (launch_haralick_order, from the
// entries' ng values, contiguous), the dispatch order of the Hessenberg and
// Schur kernels
void launch_haralick_order(const int *ng, int nimg, int *order, hipStream_t s);
void launch_haralick_hess(const HarDesc *hd, const int *order, int nimg, hipStream_t s);
void launch_haralick_qr(const HarDesc *hd, const int *order, int nimg, hipStream_t s);
void launch_haralick_final(const HarDesc *hd, OutRef o, int nimg, hipStream_t s);
void launch_tamura(const Img *imgs, const Stats *st, double *const *scratch, OutRef o, int nimg, int maxh, int maxw,
                   hipStream_t s);
void launch_comb(const Img *imgs, const double *iis, OutRef o, int nimg, hipStream_t s);
void launch_radon(const Img *imgs, const double *trig, double *const *proj, OutRef o, int nimg, int maxbins,
                  int maxh, int maxw, hipStream_t s);
// order-free projections with bounds, counts certified (proj: 8 rSize doubles
// per plane); radon_certified_fits says whether maxbins allows it
bool radon_certified_fits(int maxbins);
void launch_radon_certified(const Img *imgs, const double *trig, double *const *proj, OutRef o, int nimg, int maxbins,
                            double slack, hipStream_t s);

// objects / edges / gabor / zernike / chebyshev-fourier (k_objects.hip, k_spectral.hip)
struct ObjDesc {
  int *label;            // [n]
  int *area;             // [n+1] per label
  long long *xm, *ym;    // [n+1]
  uint64_t *keys;        // [2*(n+1)] sort buffers
  double *dist;          // [n+1]
  int *count;            // [1]
};
void launch_objects(const Img *imgs, const Stats *st, int invert, const ObjDesc *od, OutRef o, int nimg, int maxn,
                    hipStream_t s);
void launch_prewitt(const Img *imgs, const Img *mag, const Img *dir, int nimg, long maxn, hipStream_t s);
void launch_edge_final(const Img *mag, const Img *dir, const Stats *st2, const double *medians, OutRef o, int nimg,
                       hipStream_t s);
void launch_gabor_post(const Img *imgs, double *const *energy, OutRef o, int nimg, hipStream_t s);
// reduced-precision contractions (k_mixed.hip): precision 1 = FP32, 2 = BF16x3 on MFMA.
// Gabor: FP32 takes float kernels (k0); BF16x3 takes the hi / mid / lo bf16
// planes (k0, k1, k2); both write per-pixel error bounds and finish with the
// exact post-processing (FP64 recomputation of the undecided pixels, taps64)
// flags: 8 x 64 ints per image (per filter: the energies the last
// post-processing phase left open, the first 63 of their pixels);
// loop_tiles: one repair block per (image, filter) taking its tiles in turn,
// for bounds that leave few energies open
void launch_gabor_mixed(int precision, const Img *imgs, const void *k0, const void *k1, const void *k2,
                        const double *wsum, const double *taps64, double *const *energy, float *const *bound,
                        double *params, int *flags, OutRef o, int nimg, int maxh, int maxw, hipStream_t s);
void launch_gabor_post_exact(const Img *imgs, double *const *energy, float *const *bound, const double *taps,
                             double *params, int *flags, bool loop_tiles, OutRef o, int nimg, int maxh, int maxw,
                             hipStream_t s);
void launch_gabor(const Img *imgs, const double *kernels, double *const *energy, OutRef o, int nimg, int maxh,
                  int maxw, hipStream_t s);
// the plain form's default: FP64 fused multiply-adds with per-energy bounds
// (scaled by slack >= 1), every count certified by the exact post-processing
void launch_gabor_certified(const Img *imgs, const double *kernels, const double *wsum, double slack,
                            double *const *energy, float *const *bound, double *params, int *flags, OutRef o,
                            int nimg, int maxh, int maxw, hipStream_t s);
void launch_zernike(const Img *imgs, double *partial, int nchunk, const double *htab, OutRef o, int nimg,
                    hipStream_t s, int precision = 0);
// Chebyshev-Fourier histograms: matrix-core sums certified against an
// a-priori bound (slack >= 1 scales it), entries left open (open[k] = 1)
// recomputed in the reference's order
void launch_cf(const CfMGroup *mgroups, int nmgroups, int nsplit, const CfGroup *groups, int ngroups,
               const CfDesc *cd, int nimg, int *open, double slack, OutRef o, hipStream_t s);
void launch_downsample(const DsDesc *dd, int nimg, int maxoh, int maxow, int maxnh, int maxnw, hipStream_t s);
void launch_downsample_hsv(const HsvDsDesc *dd, int nimg, int maxoh, int maxow, int maxnh, int maxnw, hipStream_t s);
void launch_normalize(NormDesc *nd, int nimg, long maxn, hipStream_t s);
void launch_sample(const SampleDesc *sd, const Img *dst, int nsamples, long maxn, hipStream_t s);

// differentiable form (k_adjoint.hip): batches of n planes [n][h][w] of one size
void launch_ad_fourier_fwd(const double *in, double *out, double2 *buf, int n, int h, int w, const FftPlan &pw,
                           const FftPlan &ph, hipStream_t s);
void launch_ad_fourier_bwd(const double *in, const double *g, double *grad, double2 *buf, int n, int h, int w,
                           const FftPlan &pw, const FftPlan &ph, hipStream_t s);
void launch_ad_gemm(const AdGemm &d, int nbatch, hipStream_t s);
size_t zernike_ad_work_doubles(int n, int h, int w);
void launch_ad_zernike_fwd(const double *in, double *out, double *work, int n, int h, int w, hipStream_t s);
void launch_ad_zernike_bwd(const double *in, const double *gout, double *grad, double *work, int n, int h, int w,
                           hipStream_t s);
void launch_fill_int(int *p, long n, int v, hipStream_t s);
int cf_ad_chunks(int np);
void launch_ad_cf_slots(const int *kk, int np, long plane, int *slot, hipStream_t s);
void launch_ad_cf_fwd(const double *v, int n, long plane, const int *kk, const double *tn, const double *fc,
                      const double *fs, int np, int NN, int nchunk, double *part, double *out, hipStream_t s);
void launch_ad_cf_bwd(const double *v, const double *gout, int n, long plane, const int *kk, const int *slot,
                      const double *tn, const double *fc, const double *fs, int np, int NN, int nchunk,
                      double *part, double *coef, double *grad, hipStream_t s);
void launch_ad_comb_fwd(const double *in, double *out, int n, int h, int w, hipStream_t s);
void launch_ad_comb_bwd(const double *in, const double *gout, double *grad, double *coef, int n, int h, int w,
                        hipStream_t s);
void launch_ad_gabor_fwd(const double *in, const double *taps8, double *out, int n, int h, int w, hipStream_t s);
void launch_ad_gabor_bwd(const double *in, const double *taps8, const double *gE, double *grad, double *Z, int n,
                         int h, int w, hipStream_t s);
void launch_ad_wnd_bwd(const double *train, const int *cls, const double *wts, const double *x, const double *v,
                       const double *mins, const double *maxs, const double *dist, const double *prob,
                       const double *norm, const double *gp, double *gsn, double *gd, int n, int f, int C, int ntest,
                       double *gx, hipStream_t s);

}  // namespace wnd
