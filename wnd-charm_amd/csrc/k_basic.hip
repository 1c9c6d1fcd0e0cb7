// Memory-bound kernels of the feature path: input conversion, plane
// statistics, the Edge / Symlet-5 wavelet / Chebyshev / Fourier transforms,
// per-image radix sort (median, Gini), multiscale histograms and fractal
// features.  All FP64; operations whose order fixes integer-valued outputs
// (wavelet taps, Chebyshev contractions) follow the reference order exactly.
#include "common.hpp"
#include "color.hpp"
#include "kernels.hpp"
#include "blocksort.hpp"
#include <algorithm>
#include "fft.hpp"

namespace wnd {

// ---------------------------------------------------------------------------
// input conversion: uint16 / f32 / f64 / uint8 / RGB source with row stride ->
// FP64 plane (and the colour plane of an RGB image when the plan reads it)
__global__ void k_load(const LoadDesc *d, Img *dst) {
  const LoadDesc L = d[blockIdx.y];
  const Img o = dst[blockIdx.y];
  const long n = (long)L.h * L.w;
  double lo = 0, scale = 0;
  if (L.dtype == 5) {  // a constant image maps to 0 (the reference divides by zero there)
    lo = L.range[0];
    scale = L.range[1] > lo ? (255.0 / (L.range[1] - lo)) : 0.0;
  }
  if (L.dtype == 0 && L.stride == L.w && ((uintptr_t)L.src & 7) == 0 && ((uintptr_t)o.p & 15) == 0) {
    // contiguous uint16 (the common case): four pixels per thread, one 8-byte
    // load and two 16-byte stores
    typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));
    typedef double f64x2 __attribute__((ext_vector_type(2)));
    const gptr<const u16x4> src = (gptr<const u16x4>)to_global((const uint16_t *)L.src);
    const gptr<f64x2> dp = (gptr<f64x2>)to_global(o.p);
    const long n4 = n >> 2;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n4; i += (long)gridDim.x * blockDim.x) {
      const u16x4 v = src[i];
      dp[2 * i] = f64x2{(double)v.x, (double)v.y};
      dp[2 * i + 1] = f64x2{(double)v.z, (double)v.w};
    }
    if (blockIdx.x == 0 && threadIdx.x < (n & 3)) {
      const long i = (n4 << 2) + threadIdx.x;
      o.p[i] = (double)((const uint16_t *)L.src)[i];
    }
    return;
  }
  // planes hold fewer than 2^31 pixels: 32-bit index arithmetic
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int y = (int)i / L.w, x = (int)i - y * L.w;
    const long e = (long)y * L.stride + x;
    double v;
    if (L.dtype == 0) v = (double)((const uint16_t *)L.src)[e];
    else if (L.dtype == 1) v = (double)((const float *)L.src)[e];
    else if (L.dtype == 2) v = ((const double *)L.src)[e];
    else if (L.dtype == 3) v = (double)((const uint8_t *)L.src)[e];
    else {  // RGB (LoadTIFF with 3 samples per pixel, cmatrix.cpp:134-193): gray plane + colour plane
      unsigned char r, g, b;
      if (L.dtype == 4) {
        const uint8_t *p = (const uint8_t *)L.src + 3 * e;
        r = p[0];
        g = p[1];
        b = p[2];
      } else {
        const uint16_t *p = (const uint16_t *)L.src + 3 * e;
        r = (unsigned char)(((double)p[0] - lo) * scale);
        g = (unsigned char)(((double)p[1] - lo) * scale);
        b = (unsigned char)(((double)p[2] - lo) * scale);
      }
      v = rgb_gray(r, g, b);
      if (L.hsv) {
        const HSVb c = rgb_to_hsv(r, g, b);
        L.hsv[3 * i] = c.h;
        L.hsv[3 * i + 1] = c.s;
        L.hsv[3 * i + 2] = c.v;
      }
    }
    o.p[i] = v;
  }
}

// ---------------------------------------------------------------------------
// plane statistics (ImageMatrix::GetStats, src/cmatrix.cpp:815-827; Moments2,
// src/statistics/Moments.h:68-107).  Two-pass mean / M2 per block of 1024.
__global__ void __launch_bounds__(1024) k_stats(const Img *imgs, Stats *st) {
  __shared__ double shd[16];
  __shared__ long long shl[16];
  const Img im = imgs[blockIdx.x];
  const long n = (long)im.h * im.w;
  double s = 0, mn = 1e308, mx = -1e308;
  long long c = 0;
  for (long i = threadIdx.x; i < n; i += blockDim.x) {
    double v = im.p[i];
    if (moments_ok(v)) { s += v; c++; mn = fmin(mn, v); mx = fmax(mx, v); }
  }
  s = block_reduce(s, SumOp(), shd);
  c = block_reduce(c, SumOp(), shl);
  mn = block_reduce(mn, MinOp(), shd);
  mx = block_reduce(mx, MaxOp(), shd);
  const double mean = c ? s / (double)c : 0.0;
  double m2 = 0;
  for (long i = threadIdx.x; i < n; i += blockDim.x) {
    double v = im.p[i];
    if (moments_ok(v)) { double d = v - mean; m2 += d * d; }
  }
  m2 = block_reduce(m2, SumOp(), shd);
  if (threadIdx.x == 0) {
    Stats r;
    r.n = c;
    r.mean = mean;
    r.mn = c ? mn : 1.79769313486231570815e+308;
    r.mx = c ? mx : -1.79769313486231570815e+308;
    r.var = c > 2 ? m2 / (double)(c - 1) : 0.0;
    r.stdv = c > 2 ? sqrt(m2 / (double)(c - 1)) : 0.0;
    st[blockIdx.x] = r;
  }
}

// ---------------------------------------------------------------------------
// Edge transform (ImageMatrix::EdgeTransform, src/cmatrix.cpp:1294-1311)
__global__ void k_edge(const Img *in, const Img *out) {
  const Img a = in[blockIdx.y], o = out[blockIdx.y];
  // planes hold fewer than 2^31 pixels: 32-bit index arithmetic, global address space
  const gptr<const double> p = to_global((const double *)a.p);
  const gptr<double> q = to_global(o.p);
  const int stride = gridDim.x * blockDim.x;
  if ((a.w & 3) == 0 && (((uintptr_t)a.p | (uintptr_t)o.p) & 15) == 0) {
    // rows of a multiple of 4 pixels: four adjacent pixels per thread, centre
    // and vertical neighbours as 16-byte pairs, the two outer horizontal
    // neighbours as single loads
    typedef double f64x2 __attribute__((ext_vector_type(2)));
    const int qw = a.w >> 2, nq = a.h * qw;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nq; t += stride) {
      const int y = t / qw, x0 = (t - y * qw) << 2, i = y * a.w + x0;
      const gptr<const f64x2> pc = (gptr<const f64x2>)(p + i);
      const f64x2 c01 = pc[0], c23 = pc[1];
      const double c[4] = {c01.x, c01.y, c23.x, c23.y};
      double my[4] = {0, 0, 0, 0};
      if (y > 0 && y < a.h - 1) {
        const gptr<const f64x2> pu = (gptr<const f64x2>)(p + i - a.w), pd = (gptr<const f64x2>)(p + i + a.w);
        const f64x2 u01 = pu[0], u23 = pu[1], d01 = pd[0], d23 = pd[1];
        const double u[4] = {u01.x, u01.y, u23.x, u23.y}, d[4] = {d01.x, d01.y, d23.x, d23.y};
#pragma unroll
        for (int k = 0; k < 4; ++k) my[k] = fmax(fabs(c[k] - u[k]), fabs(c[k] - d[k]));
      }
      const double l = x0 > 0 ? p[i - 1] : 0.0, r = x0 + 4 < a.w ? p[i + 4] : 0.0;
      double e[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int x = x0 + k;
        double mx = 0;
        if (x > 0 && x < a.w - 1) {
          const double left = k > 0 ? c[k - 1] : l, right = k < 3 ? c[k + 1] : r;
          mx = fmax(fabs(c[k] - left), fabs(c[k] - right));
        }
        e[k] = fmax(mx, my[k]);
      }
      const gptr<f64x2> po = (gptr<f64x2>)(q + i);
      po[0] = f64x2{e[0], e[1]};
      po[1] = f64x2{e[2], e[3]};
    }
    return;
  }
  const int n = a.h * a.w;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int y = i / a.w, x = i - y * a.w;
    const double c = p[i];
    double my = 0, mx = 0;
    if (y > 0 && y < a.h - 1) my = fmax(fabs(c - p[i - a.w]), fabs(c - p[i + a.w]));
    if (x > 0 && x < a.w - 1) mx = fmax(fabs(c - p[i - 1]), fabs(c - p[i + 1]));
    q[i] = fmax(mx, my);
  }
}

// ---------------------------------------------------------------------------
// Symlet-5 1-level DWT, zero padding (Symlet5.cpp:10-59, Wavelet.cpp:136-434,
// convolution.cpp:86-326).  Output o of a length-N line is
//   sum_{j=max(0,i-N+1)}^{min(9,i)} in[i-j]*f[j],  i = 2o+1,
// accumulated in ascending j from 0.0 (the reference's three loop regions).
__constant__ double c_sym5_lo[10] = {0.027333068345077982, 0.029519490925774643, -0.039134249302383094,
  0.1993975339773936, 0.72340769040242059, 0.63397896345821192, 0.016602105764522319,
  -0.17532808990845047, -0.021101834024758855, 0.019538882735286728};
__constant__ double c_sym5_hi[10] = {-0.019538882735286728, -0.021101834024758855, 0.17532808990845047,
  0.016602105764522319, -0.63397896345821192, 0.72340769040242059, -0.1993975339773936,
  -0.039134249302383094, -0.029519490925774643, 0.027333068345077982};

// pass 1 (along y): g1/g2[oy][x] = approx/detail of column x
__global__ void k_wavelet_cols(const Img *in, double *const *g, const int *gstride) {
  const Img a = in[blockIdx.y];
  const int oy = (a.h + 9) / 2;
  const long n = (long)oy * a.w;
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int o = t / a.w, x = t - (long)o * a.w;
  const int i = 2 * o + 1, N = a.h;
  const int jlo = i - (N - 1) > 0 ? i - (N - 1) : 0, jhi = i < 9 ? i : 9;
  double s1 = 0, s2 = 0;
  for (int j = jlo; j <= jhi; ++j) {
    const double v = a.p[(long)(i - j) * a.w + x];
    s1 += v * c_sym5_lo[j];
    s2 += v * c_sym5_hi[j];
  }
  double *g1 = g[blockIdx.y];
  g1[t] = s1;
  g1[(long)gstride[blockIdx.y] + t] = s2;
}

// pass 2 (along x) + quadrant placement: with OX = floor((W+9)/2),
// OY = floor((H+9)/2), the output is 2*OY x 2*OX with
//   (y<OY, x<OX) = dd, (y>=OY, x<OX) = da, (y<OY, x>=OX) = ad, (y>=OY, x>=OX) = aa
// (first letter: along y, second: along x; Wavelet.cpp:348-428, nsteps = 1).
__global__ void k_wavelet_rows(const Img *in, double *const *g, const int *gstride, const Img *out) {
  const Img a = in[blockIdx.y], o = out[blockIdx.y];
  const int OY = (a.h + 9) / 2, OX = (a.w + 9) / 2, N = a.w;
  const long n = (long)OY * OX;
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int r = t / OX, ox = t - (long)r * OX;
  const double *g1 = g[blockIdx.y] + (long)r * N;
  const double *g2 = g1 + gstride[blockIdx.y];
  const int i = 2 * ox + 1;
  const int jlo = i - (N - 1) > 0 ? i - (N - 1) : 0, jhi = i < 9 ? i : 9;
  double a1 = 0, d1 = 0, a2 = 0, d2 = 0;
  for (int j = jlo; j <= jhi; ++j) {
    const double v1 = g1[i - j], v2 = g2[i - j];
    a1 += v1 * c_sym5_lo[j];
    d1 += v1 * c_sym5_hi[j];
    a2 += v2 * c_sym5_lo[j];
    d2 += v2 * c_sym5_hi[j];
  }
  const int W = 2 * OX;
  o.p[(long)r * W + ox] = d2;
  o.p[(long)(r + OY) * W + ox] = a2;
  o.p[(long)r * W + ox + OX] = d1;
  o.p[(long)(r + OY) * W + ox + OX] = a1;
}

// ---------------------------------------------------------------------------
// Chebyshev contraction (Chebyshev2D / getChCoeff1D, src/transforms/chebyshev.cpp:80-165):
//   C[r][c] = sum_{a<K} (A(r,a) * B[a][c]) / 2, accumulated in ascending a
// so the result is bit-identical to the reference's loop.  B holds the
// host-computed basis tj = T/W (c = 0) or T*2/W.  64x64 output tile per block,
// 4x4 outputs per thread, K staged through LDS 16 at a time.
__global__ void __launch_bounds__(256) k_cheb_gemm(const GemmDesc *ds) {
  const GemmDesc d = ds[blockIdx.z];
  const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  if (r0 >= d.M || c0 >= d.N) return;
  __shared__ double As[16][65];
  __shared__ double Bs[16][65];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = 0.0;
  for (int k0 = 0; k0 < d.K; k0 += 16) {
    int odd = 0;  // a nonzero operand outside [2^-500, 2^500] (or not finite)
    for (int e = threadIdx.x; e < 16 * 64; e += 256) {
      const int kk = e >> 6, rr = e & 63;
      const int k = k0 + kk, r = r0 + rr, c = c0 + rr;
      const double a = (k < d.K && r < d.M) ? d.A[(long)r * d.sar + (long)k * d.sak] : 0.0;
      const double b = (k < d.K && c < d.N) ? d.B[(long)k * d.N + c] : 0.0;
      As[kk][rr] = a;
      Bs[kk][rr] = b;
      odd |= (a != 0.0 && !(fabs(a) >= 0x1p-500 && fabs(a) <= 0x1p500)) ||
             (b != 0.0 && !(fabs(b) >= 0x1p-500 && fabs(b) <= 0x1p500));
    }
    // With every nonzero operand in [2^-500, 2^500], (a / 2) b rounds to the
    // same value as (a b) / 2: both are the product's correctly rounded value
    // scaled by 2^-1, no step underflows, and a / 2 is exact.  A tile with any
    // other operand takes the reference's expression.
    const bool scaled = !__syncthreads_or(odd);
    const int kn = d.K - k0 < 16 ? d.K - k0 : 16;
    if (scaled) {
      for (int kk = 0; kk < kn; kk++) {
        double av[4], bv[4];
#pragma unroll
        for (int i = 0; i < 4; i++) av[i] = As[kk][ty * 4 + i] * 0.5;
#pragma unroll
        for (int j = 0; j < 4; j++) bv[j] = Bs[kk][tx * 4 + j];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int j = 0; j < 4; j++) acc[i][j] += av[i] * bv[j];
      }
    } else {
      for (int kk = 0; kk < kn; kk++) {
        double av[4], bv[4];
#pragma unroll
        for (int i = 0; i < 4; i++) av[i] = As[kk][ty * 4 + i];
#pragma unroll
        for (int j = 0; j < 4; j++) bv[j] = Bs[kk][tx * 4 + j];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int j = 0; j < 4; j++) acc[i][j] += (av[i] * bv[j]) / 2;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int r = r0 + ty * 4 + i, c = c0 + tx * 4 + j;
      if (r < d.M && c < d.N) d.C[(long)r * d.N + c] = acc[i][j];
    }
}

// ---------------------------------------------------------------------------
// Fourier magnitude (ImageMatrix::fft2, src/cmatrix.cpp:1107-1160): the
// unnormalised 2-D DFT, |F(ky,kx)| for ky < H/2+1, the lower rows filled by
// |F(H-ky, (W-kx) mod W)| as the reference does.  Mixed-radix Stockham FFT
// in LDS (fft.hpp), one work-group per line.

// rows: F1[y][x] = DFT_x(img[y][:]), complex, row-major H x W; fft_lines(W)
// rows per block, each on its own group of threads and its own LDS pair
__global__ void __launch_bounds__(256) k_fft_rows(const Img *in, double2 *const *f1, const FftPlan *plw) {
  extern __shared__ double2 sh[];
  const Img a = in[blockIdx.y];
  const FftPlan &pl = plw[blockIdx.y];
  const int L = fft_lines(pl.n), tpl = blockDim.x / L, line = threadIdx.x / tpl, lane = threadIdx.x - line * tpl;
  const int row = blockIdx.x * L + line;
  if ((int)blockIdx.x * L >= a.h) return;
  double2 *b0 = sh + (long)line * 2 * pl.n, *b1 = b0 + pl.n;
  for (int x = lane; x < a.w; x += tpl) b0[x] = make_double2(row < a.h ? a.p[(long)row * a.w + x] : 0.0, 0.0);
  double2 *twl = sh + (long)L * 2 * pl.n;  // twiddles staged after the lines (n <= 1024)
  if (pl.n <= 1024)
    for (int k = threadIdx.x; k < pl.n; k += blockDim.x) twl[k] = pl.tw[k];
  __syncthreads();
  double2 *r;
  if (pl.n <= 1024) fft_lds_group_tw(b0, b1, pl, (const double2 *)twl, lane, tpl, &r);
  else fft_lds_group(b0, b1, pl, lane, tpl, &r);
  if (row < a.h) {
    double2 *o = f1[blockIdx.y] + (long)row * a.w;
    for (int x = lane; x < a.w; x += tpl) o[x] = r[x];
  }
}

// columns: out(ky, x) = |DFT_y(F1[:][x])| for ky < H/2+1; fft_lines(H)
// adjacent columns per block, loaded and stored row by row (coalesced)
__global__ void __launch_bounds__(256) k_fft_cols(const Img *in, double2 *const *f1, const FftPlan *plh,
                                                  const Img *out) {
  extern __shared__ double2 sh[];
  const Img a = in[blockIdx.y], o = out[blockIdx.y];
  const FftPlan &pl = plh[blockIdx.y];
  const int n = pl.n, L = fft_lines(n), tpl = blockDim.x / L, line = threadIdx.x / tpl, lane = threadIdx.x - line * tpl;
  const int x0 = blockIdx.x * L;
  if (x0 >= a.w) return;
  const double2 *src = f1[blockIdx.y];
  for (int e = threadIdx.x; e < a.h * L; e += blockDim.x) {
    const int y = e / L, l = e - y * L, x = x0 + l;
    sh[(long)l * 2 * n + y] = x < a.w ? src[(long)y * a.w + x] : make_double2(0.0, 0.0);
  }
  double2 *twl = sh + (long)L * 2 * n;  // twiddles staged after the lines (n <= 1024)
  if (n <= 1024)
    for (int k = threadIdx.x; k < n; k += blockDim.x) twl[k] = pl.tw[k];
  __syncthreads();
  double2 *b0 = sh + (long)line * 2 * n, *b1 = b0 + n;
  double2 *r;
  if (n <= 1024) fft_lds_group_tw(b0, b1, pl, (const double2 *)twl, lane, tpl, &r);
  else fft_lds_group(b0, b1, pl, lane, tpl, &r);
  const long roff = (r == b1) ? n : 0;  // the same buffer of the pair for every line
  const int hh = a.h / 2 + 1;
  for (int e = threadIdx.x; e < hh * L; e += blockDim.x) {
    const int y = e / L, l = e - y * L, x = x0 + l;
    if (x < a.w) {
      const double2 c = sh[(long)l * 2 * n + roff + y];
      o.p[(long)y * a.w + x] = sqrt(c.x * c.x + c.y * c.y);
    }
  }
}

__global__ void k_fft_fill(const Img *out) {
  const Img o = out[blockIdx.y];
  const int hh = o.h / 2 + 1;
  const long n = (long)(o.h - hh) * o.w;
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int y = hh + t / o.w, x = t % o.w;
  const int xs = x == 0 ? 0 : o.w - x;
  o.p[(long)y * o.w + x] = o.p[(long)(o.h - y) * o.w + xs];
}

// ---------------------------------------------------------------------------
// Per-image radix sort of the plane (median, Gini); one 1024-thread
// work-group per image, result left in buf[0..n).
__global__ void __launch_bounds__(1024) k_sort(const Img *imgs, uint64_t *const *buf, const long *cap) {
  const Img im = imgs[blockIdx.x];
  const long n = (long)im.h * im.w;
  uint64_t *k0 = buf[blockIdx.x], *k1 = k0 + cap[blockIdx.x];
  for (long i = threadIdx.x; i < n; i += 1024) k0[i] = dkey(im.p[i]);
  __syncthreads();
  uint64_t *r = block_radix_sort(k0, k1, n);
  if (r != k0)
    for (long i = threadIdx.x; i < n; i += 1024) k0[i] = r[i];
}

// median (ImageMatrix::get_median, cmatrix.cpp:780-800) and Gini
// (FeatureAlgorithms.cpp:448-489) from the sorted keys; writes the five
// Pixel Intensity Statistics (FeatureAlgorithms.cpp:272-291) and/or Gini.
__global__ void __launch_bounds__(1024) k_median_gini(const Img *imgs, uint64_t *const *buf, const Stats *st,
                                                      double *median_out, OutRef pix, OutRef gini) {
  __shared__ double shd[16];
  __shared__ long long shl[16];
  __shared__ long long first_pos;
  const Img im = imgs[blockIdx.x];
  const long n = (long)im.h * im.w;
  const uint64_t *k = buf[blockIdx.x];
  if (threadIdx.x == 0) first_pos = n;
  __syncthreads();
  // first index whose value is > 0 (keys are sorted)
  for (long i = threadIdx.x; i < n; i += 1024) {
    const double v = dkey_inv(k[i]);
    const double pv = i ? dkey_inv(k[i - 1]) : -1.0;
    if (v > 0 && !(pv > 0)) atomicMin((unsigned long long *)&first_pos, (unsigned long long)i);
  }
  __syncthreads();
  const long half = n / 2;
  double med;
  if (n % 2 == 0) { med = dkey_inv(k[half]); med += dkey_inv(k[half - 1]); med /= 2.0; }
  else med = dkey_inv(k[half]);
  const long fp = first_pos, cnt = n - fp;
  double s = 0, g = 0;
  for (long i = fp + threadIdx.x; i < n; i += 1024) {
    const double v = dkey_inv(k[i]);
    const long ii = i - fp + 1;
    s += v;
    g += (2. * ii - cnt - 1.) * v;
  }
  s = block_reduce(s, SumOp(), shd);
  g = block_reduce(g, SumOp(), shd);
  if (threadIdx.x == 0) {
    if (median_out) median_out[blockIdx.x] = med;
    if (pix.rows && pix.rows[blockIdx.x]) {
      double *o = pix.rows[blockIdx.x];
      const Stats S = st[blockIdx.x];
      o[0] = S.mean; o[1] = med; o[2] = S.stdv; o[3] = S.mn; o[4] = S.mx;
    }
    if (gini.rows && gini.rows[blockIdx.x]) {
      const double mean = cnt > 0 ? s / cnt : 0.0;
      double r;
      if (cnt <= 1 || mean <= 0.0) r = 0.0;
      else r = g / (mean * cnt * (cnt - 1));
      gini.rows[blockIdx.x][0] = r;
    }
  }
}

// ---------------------------------------------------------------------------
// Multiscale histograms (ImageMatrix::MultiScaleHistogram, cmatrix.cpp:1743-1756)
__global__ void __launch_bounds__(1024) k_msh(const Img *imgs, const Stats *st, OutRef o) {
  __shared__ int cnt[24];
  const Img im = imgs[blockIdx.x];
  const Stats S = st[blockIdx.x];
  if (threadIdx.x < 24) cnt[threadIdx.x] = 0;
  __syncthreads();
  const long n = (long)im.h * im.w;
  const double s3 = hist_scale(S.mn, S.mx, 3), s5 = hist_scale(S.mn, S.mx, 5), s7 = hist_scale(S.mn, S.mx, 7),
               s9 = hist_scale(S.mn, S.mx, 9);
  // each wave counts its 64 pixels per bin by ballot (wave-uniform counts,
  // kept in scalar registers) and adds its 24 totals into LDS once at the end
  int wc[24];
#pragma unroll
  for (int b = 0; b < 24; b++) wc[b] = 0;
  for (long i0 = 0; i0 < n; i0 += blockDim.x) {
    const long i = i0 + threadIdx.x;
    const bool ok = i < n;
    const double v = ok ? im.p[i] : S.mn;
    const int b3 = hist_bin(v, S.mn, s3, 3), b5 = hist_bin(v, S.mn, s5, 5), b7 = hist_bin(v, S.mn, s7, 7),
              b9 = hist_bin(v, S.mn, s9, 9);
#pragma unroll
    for (int b = 0; b < 3; b++) wc[b] += __popcll(__ballot(ok && b3 == b));
#pragma unroll
    for (int b = 0; b < 5; b++) wc[3 + b] += __popcll(__ballot(ok && b5 == b));
#pragma unroll
    for (int b = 0; b < 7; b++) wc[8 + b] += __popcll(__ballot(ok && b7 == b));
#pragma unroll
    for (int b = 0; b < 9; b++) wc[15 + b] += __popcll(__ballot(ok && b9 == b));
  }
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int b = 0; b < 24; b++) atomicAdd(&cnt[b], wc[b]);
  __syncthreads();
  if (threadIdx.x == 0) {
    double mx = 0;
    for (int a = 0; a < 24; a++) if (cnt[a] > mx) mx = cnt[a];
    double *out = o.rows[blockIdx.x];
    for (int a = 0; a < 24; a++) out[a] = (double)cnt[a] / mx;
  }
}

// ---------------------------------------------------------------------------
// Fractal features (FeatureAlgorithms.cpp:232-264): one block per (k, image)
__global__ void __launch_bounds__(256) k_fractal(const Img *imgs, OutRef o, int nimg) {
  __shared__ double shd[4];
  int ent, bin;
  if (!xcd_item(20, nimg, ent, bin)) return;  // the 20 offsets of one plane share an XCD
  const Img im = imgs[ent];
  const int W = im.w, H = im.h;
  const int K = (W > H ? H : W) / 5;
  int step = K / 20;
  if (step < 1) step = 1;
  const int k = 1 + bin * step;
  double *out = o.rows[ent];
  if (k >= K) {
    if (threadIdx.x == 0) out[bin] = 0.0;
    return;
  }
  // sum |I(y,x) - I(y+k,x)| + |I(y,x) - I(y,x+k)| in one pass: rows by
  // 4-row groups of 64-lane row segments
  double s = 0;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int y = wv; y < H; y += 4) {
    const double *row = im.p + (long)y * W;
    const double *down = row + (long)k * W;
    const bool hasdown = y + k < H;
    for (int x = lane; x < W; x += 64) {
      const double v = row[x];
      if (hasdown) s += fabs(v - down[x]);
      if (x + k < W) s += fabs(v - row[x + k]);
    }
  }
  s = block_reduce(s, SumOp(), shd);
  if (threadIdx.x == 0) out[bin] = s / (W * (W - k) + H * (H - k));
}

// The same 20 sums with one block per plane and each pixel read from global
// memory about once: the block takes 64-column bands in turn and walks their
// rows in chunks of FRACTAL_CH, keeping rows y .. y + kmax + FRACTAL_CH - 1
// of the band (plus kmax columns on its right) in an LDS ring, so the
// vertical and horizontal differences of every row of a chunk at every offset
// come from LDS; the next chunk's rows are loaded (one column per thread)
// while a chunk is computed, with two barriers per chunk.
// Thread t takes column t & 63 of the band and the offsets t >> 6, +4, +8,
// +12, +16; the per-offset partial sums are added in a fixed order at the
// end.  The ring has a power-of-two number of rows (slot = row & (R - 1)).
// Used when it fits FRACTAL_LDS bytes.
constexpr int FRACTAL_LDS = 98304, FRACTAL_CH = 16;
__host__ __device__ inline int fractal_ring_rows(int kmax) {
  int r = 16;
  while (r < kmax + FRACTAL_CH) r *= 2;
  return r;
}
__device__ __forceinline__ void fractal_geometry(int W, int H, int &K, int &step, int &nk) {
  K = (W > H ? H : W) / 5;
  step = K / 20;
  if (step < 1) step = 1;
  nk = K > 1 ? min(20, (K - 2) / step + 1) : 0;  // offsets 1 + j step < K, j < 20
}
__global__ void __launch_bounds__(256) k_fractal_ring(const Img *imgs, OutRef o) {
  extern __shared__ double ring[];
  __shared__ double part[20 * 64];
  const Img im = imgs[blockIdx.x];
  const int W = im.w, H = im.h;
  int K, step, nk;
  fractal_geometry(W, H, K, step, nk);
  const int kmax = nk > 0 ? 1 + (nk - 1) * step : 0, RM = fractal_ring_rows(kmax) - 1, LW = 64 + kmax;
  const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
  const gptr<const double> P = to_global((const double *)im.p);
  double acc[5] = {0, 0, 0, 0, 0};
  for (int x0 = 0; x0 < W && nk > 0; x0 += 64) {
    const int lw = min(LW, W - x0);  // columns of the band and its right margin (lw <= 256)
    const int x = x0 + c;
    const bool ld = (int)threadIdx.x < lw;  // thread t loads column x0 + t of the band
    const gptr<const double> Pc = P + x0 + (ld ? (int)threadIdx.x : 0);
    // the first chunk's rows 0 .. FRACTAL_CH + kmax - 1
    if (ld)
      for (int yy = 0; yy < min(H, FRACTAL_CH + kmax); yy++) ring[(yy & RM) * LW + threadIdx.x] = Pc[(long)yy * W];
    __syncthreads();
    for (int y0 = 0; y0 < H; y0 += FRACTAL_CH) {
      // the next chunk's new rows nr .. nr + FRACTAL_CH - 1, loaded while this
      // chunk is computed and written once its rows are consumed (their slots
      // held rows at least R rows earlier, below y0 + FRACTAL_CH)
      const int nr = y0 + FRACTAL_CH + kmax;
      double pre[FRACTAL_CH];
#pragma unroll
      for (int j = 0; j < FRACTAL_CH; j++) pre[j] = (ld && nr + j < H) ? Pc[(long)(nr + j) * W] : 0.0;
      if (x < W) {
        for (int y = y0; y < min(H, y0 + FRACTAL_CH); y++) {
          const double *row = ring + (y & RM) * LW;
          const double v = row[c];
#pragma unroll
          for (int i = 0; i < 5; i++) {
            const int j = g + 4 * i, k = 1 + j * step;
            if (j < nk) {
              if (y + k < H) acc[i] += fabs(v - ring[((y + k) & RM) * LW + c]);
              if (x + k < W) acc[i] += fabs(v - row[c + k]);
            }
          }
        }
      }
      __syncthreads();  // the chunk's rows are consumed before their slots refill
      if (ld)
#pragma unroll
        for (int j = 0; j < FRACTAL_CH; j++)
          if (nr + j < H) ring[((nr + j) & RM) * LW + threadIdx.x] = pre[j];
      __syncthreads();
    }
  }
#pragma unroll
  for (int i = 0; i < 5; i++) part[(g + 4 * i) * 64 + c] = acc[i];
  __syncthreads();
  if (threadIdx.x < 20) {
    const int j = threadIdx.x, k = 1 + j * step;
    double sum = 0;
    for (int l = 0; l < 64; l++) sum += part[j * 64 + l];
    o.rows[blockIdx.x][j] = j < nk ? sum / (W * (W - k) + H * (H - k)) : 0.0;
  }
}

// ---------------------------------------------------------------------------
// Small dense histogram of a value list with its own min/max (used for the
// Chebyshev-20 coefficient block, ImageMatrix::ChebyshevStatistics2D,
// cmatrix.cpp:1274-1281).
__global__ void __launch_bounds__(256) k_hist_values(const Img *vals, int nbins, OutRef o) {
  __shared__ double shd[4];
  __shared__ int cnt[64];
  const Img v = vals[blockIdx.x];
  const long n = (long)v.h * v.w;
  double mn = 1e308, mx = -1e308;
  for (long i = threadIdx.x; i < n; i += blockDim.x) {
    const double x = v.p[i];
    if (moments_ok(x)) { mn = fmin(mn, x); mx = fmax(mx, x); }
  }
  mn = block_reduce(mn, MinOp(), shd);
  mx = block_reduce(mx, MaxOp(), shd);
  if (threadIdx.x < 64) cnt[threadIdx.x] = 0;
  __syncthreads();
  const double sc = hist_scale(mn, mx, nbins);
  for (long i = threadIdx.x; i < n; i += blockDim.x) atomicAdd(&cnt[hist_bin(v.p[i], mn, sc, nbins)], 1);
  __syncthreads();
  if (threadIdx.x < nbins) o.rows[blockIdx.x][threadIdx.x] = cnt[threadIdx.x];
}

// gathers per-entry statistics into one array (entries of a combined launch
// come from different DAG nodes)
__global__ void k_gather_stats(const Stats *const *src, Stats *dst, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = *src[i];
}

// ===========================================================================
// launchers
void launch_gather_stats(const Stats *const *src, Stats *dst, int n, hipStream_t s) {
  k_gather_stats<<<(n + 255) / 256, 256, 0, s>>>(src, dst, n);
}
void launch_load(const LoadDesc *d, Img *dst, int nimg, long maxn, hipStream_t s) {
  // four pixels per thread on the contiguous uint16 path; the other paths stride over the plane
  int blocks = (int)((maxn + 1023) / 1024);
  if (blocks > 4096) blocks = 4096;
  k_load<<<dim3(blocks, nimg), 256, 0, s>>>(d, dst);
}
void launch_stats(const Img *imgs, Stats *st, int nimg, hipStream_t s) {
  k_stats<<<nimg, 1024, 0, s>>>(imgs, st);
}
void launch_edge(const Img *in, const Img *out, int nimg, long maxn, hipStream_t s) {
  // four pixels per thread where rows allow it; otherwise each thread strides over the plane
  int blocks = (int)((maxn + 1023) / 1024);
  if (blocks > 4096) blocks = 4096;
  k_edge<<<dim3(blocks, nimg), 256, 0, s>>>(in, out);
}
void launch_wavelet(const Img *in, double *const *g, const int *gstride, const Img *out, int nimg, int maxh,
                    int maxw, hipStream_t s) {
  const long n1 = (long)((maxh + 9) / 2) * maxw;
  k_wavelet_cols<<<dim3((n1 + 255) / 256, nimg), 256, 0, s>>>(in, g, gstride);
  const long n2 = (long)((maxh + 9) / 2) * ((maxw + 9) / 2);
  k_wavelet_rows<<<dim3((n2 + 255) / 256, nimg), 256, 0, s>>>(in, g, gstride, out);
}
void launch_cheb_gemm(const GemmDesc *d, int nimg, int maxM, int maxN, hipStream_t s) {
  k_cheb_gemm<<<dim3((maxN + 63) / 64, (maxM + 63) / 64, nimg), 256, 0, s>>>(d);
}
void launch_fft(const Img *in, double2 *const *f1, const FftPlan *plw, const FftPlan *plh, const Img *out, int nimg,
                int maxh, int maxw, hipStream_t s) {
  const int nmax = maxh > maxw ? maxh : maxw;
  // per block: fft_lines(n) lines of 2n entries, then n twiddles for
  // n <= 1024; the largest over n <= nmax
  const int ent = nmax <= 256 ? 4 * 2 * 256 + 256 : nmax <= 512 ? 2 * 2 * 512 + 512 : nmax <= 1024 ? 2 * 1024 + 1024
                                                                                                      : std::max(2 * nmax, 3072);
  const size_t lds = (size_t)ent * sizeof(double2);
  static const bool attr = [] {
    (void)hipFuncSetAttribute((const void *)k_fft_rows, hipFuncAttributeMaxDynamicSharedMemorySize, 2 * 4096 * 16);
    (void)hipFuncSetAttribute((const void *)k_fft_cols, hipFuncAttributeMaxDynamicSharedMemorySize, 2 * 4096 * 16);
    return true;
  }();
  (void)attr;
  // every image's row (column) count fits in its blocks: fft_lines is smallest for the longest line
  const int lr = fft_lines(maxw), lc = fft_lines(maxh);
  k_fft_rows<<<dim3((maxh + lr - 1) / lr, nimg), 256, lds, s>>>(in, f1, plw);
  k_fft_cols<<<dim3((maxw + lc - 1) / lc, nimg), 256, lds, s>>>(in, f1, plh, out);
  const long nf = (long)(maxh - (maxh / 2 + 1) + 1) * maxw;
  k_fft_fill<<<dim3((nf + 255) / 256, nimg), 256, 0, s>>>(out);
}
void launch_sort(const Img *imgs, uint64_t *const *buf, const long *cap, int nimg, hipStream_t s) {
  k_sort<<<nimg, 1024, 0, s>>>(imgs, buf, cap);
}
void launch_median_gini(const Img *imgs, uint64_t *const *buf, const Stats *st, double *median_out, OutRef pix,
                        OutRef gini, int nimg, hipStream_t s) {
  k_median_gini<<<nimg, 1024, 0, s>>>(imgs, buf, st, median_out, pix, gini);
}
void launch_msh(const Img *imgs, const Stats *st, OutRef o, int nimg, hipStream_t s) {
  k_msh<<<nimg, 1024, 0, s>>>(imgs, st, o);
}
void launch_fractal(const Img *imgs, OutRef o, int nimg, int maxh, int maxw, hipStream_t s) {
  // the ring's size for the largest offset any image of the batch can have
  // (the offsets grow with the smaller side)
  const int side = std::min(maxh, maxw), K = side / 5, step = std::max(1, K / 20);
  const int nk = K > 1 ? std::min(20, (K - 2) / step + 1) : 0, kmax = nk > 0 ? 1 + (nk - 1) * step : 0;
  const size_t ring = (size_t)fractal_ring_rows(kmax) * (64 + kmax) * sizeof(double);
  if (ring <= (size_t)FRACTAL_LDS) {
    static const bool attr = hipFuncSetAttribute((const void *)k_fractal_ring,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, FRACTAL_LDS) == hipSuccess;
    (void)attr;
    k_fractal_ring<<<nimg, 256, ring, s>>>(imgs, o);
  } else {
    k_fractal<<<xcd_grid(20, nimg), 256, 0, s>>>(imgs, o, nimg);
  }
}
void launch_hist_values(const Img *vals, int nbins, OutRef o, int nimg, hipStream_t s) {
  k_hist_values<<<nimg, 256, 0, s>>>(vals, nbins, o);
}

}  // namespace wnd
