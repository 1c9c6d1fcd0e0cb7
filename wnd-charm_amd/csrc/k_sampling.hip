// Sampling of the images before the feature path: the colour plane's part of
// ImageMatrix::Downsample (src/cmatrix.cpp:579-720; the pixel plane uses the
// Chebyshev-Fourier downsample kernels), mean / stdev normalisation
// (ImageMatrix::normalize, cmatrix.cpp:836-888) and the rotated, tiled samples
// (Rotate, cmatrix.cpp:727-767; the tiles of TrainingSet::AddImageFile,
// src/wndchrm_src/TrainingSet.cpp:1455-1480).  All HBM-bound gathers.
#include "common.hpp"
#include "kernels.hpp"

namespace wnd {

// One output of a Downsample pass over the colour plane: per channel the
// reference's left fraction + whole pixels + right fraction, / d, truncated
// to a byte.  v: first pixel of the line, stride: pixels between line entries.
__device__ __forceinline__ void ds_line_hsv(const unsigned char *v, long stride, unsigned len, double x, double d,
                                            unsigned char *out) {
  double s0 = 0, s1 = 0, s2 = 0;
  unsigned a = (unsigned)(floor(x));
  double frac = ceil(x) - x;
  if (frac > 0 && a < len) {
    const unsigned char *p = v + 3 * (long)a * stride;
    s0 += p[0] * frac;
    s1 += p[1] * frac;
    s2 += p[2] * frac;
  }
  for (a = (unsigned)(ceil(x)); a < floor(x + d); a++)
    if (a < len) {
      const unsigned char *p = v + 3 * (long)a * stride;
      s0 += p[0];
      s1 += p[1];
      s2 += p[2];
    }
  frac = x + d - floor(x + d);
  if (frac > 0 && a < len) {
    const unsigned char *p = v + 3 * (long)a * stride;
    s0 += p[0] * frac;
    s1 += p[1] * frac;
    s2 += p[2] * frac;
  }
  out[0] = (unsigned char)(s0 / (d));
  out[1] = (unsigned char)(s1 / (d));
  out[2] = (unsigned char)(s2 / (d));
}

__global__ void k_ds_hsv_x(const HsvDsDesc *dds) {
  const HsvDsDesc D = dds[blockIdx.y];
  const long n = (long)D.oh * D.cx;
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int y = t / D.cx, k = t - (long)y * D.cx;
  ds_line_hsv(D.in + 3 * (long)y * D.ow, 1, D.ow, D.xs[k], D.dx, D.tmp + 3 * ((long)y * D.ow + k));
}

__global__ void k_ds_hsv_y(const HsvDsDesc *dds) {
  const HsvDsDesc D = dds[blockIdx.y];
  const long n = (long)D.nh * D.nw;
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int y = t / D.nw, x = t - (long)y * D.nw;
  ds_line_hsv(D.tmp + 3 * (long)x, D.ow, D.oh, D.ys[y], D.dy, D.out + 3 * ((long)y * D.nw + x));
}

// normalize(-1, -1, -1, mean, stdev): the plane's mean and stdev are Moments2
// (src/statistics/Moments.h:60-94) accumulated in raster order, as
// ImageMatrix::UpdateStats does; that sequential order fixes every rounding,
// so one lane per image runs it.  Then every pixel: v - (mean - n_mean),
// scaled about n_mean by n_std / std when n_std > 0, clamped to [0, max].
__global__ void k_norm_stats(NormDesc *nd) {
  NormDesc &N = nd[blockIdx.x];
  if (threadIdx.x != 0) return;
  const gptr<const double> p = to_global((const double *)N.p);
  double mean = 0, m2 = 0;
  unsigned long long n = 0;
  for (long i = 0; i < N.n; i++) {
    const double x = p[i];
    if (!moments_ok(x)) continue;
    const unsigned long long n1 = n;
    n = n + 1;
    const double delta = x - mean;
    const double delta_n = delta / (double)n;
    const double term1 = delta * delta_n * (double)n1;
    mean = mean + delta_n;
    m2 += term1;
  }
  const double sd = n > 2 ? sqrt(m2 / (double)(n - 1)) : 0.0;
  N.d_mean = mean - N.mean;
  N.std_fact = (N.stdev > 0 ? N.stdev : 0) / sd;
}

__global__ void k_norm_apply(const NormDesc *nd) {
  const NormDesc N = nd[blockIdx.y];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N.n; i += (long)gridDim.x * blockDim.x) {
    double val = N.p[i] - N.d_mean;
    if (N.stdev > 0) val = N.mean + (val - N.mean) * N.std_fact;
    if (val < 0) val = 0;
    else if (val > N.max_range) val = N.max_range;
    N.p[i] = val;
  }
}

// sample (i, j) = pixel (y0 + i, x0 + j) of the source rotated by 90 * rot
// degrees: 90 is transpose().rowwise().reverse(), 180 reverse(), and the
// reference's 270 is transpose()
__global__ void k_sample(const SampleDesc *sd, const Img *dst) {
  const SampleDesc S = sd[blockIdx.y];
  const Img o = dst[blockIdx.y];
  const long n = (long)o.h * o.w;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const int i = (int)(t / o.w) + S.y0, j = (int)(t % o.w) + S.x0;
    long src;
    if (S.rot == 1) src = (long)(S.sh - 1 - j) * S.sw + i;
    else if (S.rot == 2) src = (long)(S.sh - 1 - i) * S.sw + (S.sw - 1 - j);
    else if (S.rot == 3) src = (long)j * S.sw + i;
    else src = (long)i * S.sw + j;
    o.p[t] = S.src[src];
    if (S.hdst) {
      S.hdst[3 * t] = S.hsrc[3 * src];
      S.hdst[3 * t + 1] = S.hsrc[3 * src + 1];
      S.hdst[3 * t + 2] = S.hsrc[3 * src + 2];
    }
  }
}

void launch_downsample_hsv(const HsvDsDesc *dd, int nimg, int maxoh, int maxow, int maxnh, int maxnw, hipStream_t s) {
  const long n1 = (long)maxoh * maxow, n2 = (long)maxnh * maxnw;
  k_ds_hsv_x<<<dim3((n1 + 255) / 256, nimg), 256, 0, s>>>(dd);
  k_ds_hsv_y<<<dim3((n2 + 255) / 256, nimg), 256, 0, s>>>(dd);
}
void launch_normalize(NormDesc *nd, int nimg, long maxn, hipStream_t s) {
  k_norm_stats<<<nimg, 64, 0, s>>>(nd);
  long blocks = (maxn + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  k_norm_apply<<<dim3(blocks, nimg), 256, 0, s>>>(nd);
}
void launch_sample(const SampleDesc *sd, const Img *dst, int nsamples, long maxn, hipStream_t s) {
  long blocks = (maxn + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  k_sample<<<dim3(blocks, nsamples), 256, 0, s>>>(sd, dst);
}

}  // namespace wnd
