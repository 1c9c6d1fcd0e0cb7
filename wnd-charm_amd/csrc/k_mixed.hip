// Reduced-precision forms of the Gabor filter bank (GaborEnergy's
// conv2comp, src/textures/gabor.cpp:19-52, 166-195) for the mixed-precision
// plans:
//   * FP32: the FP64 kernel's tiling with FP32 fused multiply-adds;
//   * BF16: pixels split into two bf16 parts and weights into three, five
//     exact bf16 products per tap on the bf16 matrix cores
//     (v_mfma_f32_16x16x16_bf16), FP32 chunks added to FP64 sums.
// Each energy comes with an a-priori bound on its distance from the FP64
// kernel's value; the FP64 post-processing certifies every count from those
// bounds and recomputes in FP64 only the energies a count leaves open.
#include "common.hpp"
#include "kernels.hpp"

namespace wnd {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bf16, round to nearest even (finite inputs)
__device__ __forceinline__ unsigned bf16_bits(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}
// hi | lo << 16 with f ~= hi + lo
__device__ __forceinline__ unsigned bf16_split(float f) {
  const unsigned hi = bf16_bits(f);
  const float rest = f - __uint_as_float(hi << 16);
  return hi | (bf16_bits(rest) << 16);
}

// ---------------------------------------------------------------- Gabor
// Both reduced forms write, beside each energy E, an a-priori bound D on
// |E - E64|, E64 being the FP64 kernel's value (k_gabor_conv_t<false>), so that the
// exact post-processing (k_gabor_post_exact) can certify every count:
//   re, im = sum_{j,i} x w over the 38 x 38 taps, S = sum |x| (|wr| + |wi|)
//   <= maxwin(|x|) * wsum_f (maxwin: the largest |x| in the pixel's window,
//   wsum_f = sum of |wr| + |wi| of filter f), |E - E64| <= |dre| + |dim| +
//   the rounding of the two FP64 magnitudes.  Each kernel row (FP32) or
//   MFMA chunk (BF16) is accumulated in FP32 from zero and added to FP64
//   running sums, so the FP32 error is that of one short sum:
//   FP32   groups of 8 taps of a row, fused multiply-adds from zero:
//          gamma_8 + 2u (x, w to FP32) -> D = 11 u S  (u = 2^-24)
//   BF16   x = xh + xl exactly for integers below 2^16 (else the tile's
//          largest relative residual eps_x); w = wh + wm + wl to u; five
//          exact bf16 products xh wh, xh wm, xh wl, xl wh, xl wm (xl wl
//          <= u |x w| dropped), 80 products per chunk in FP32
//          -> D = (85 u + eps_x) S
//   plus 2e-13 S for the FP64 kernel's own summation and 1e-14 E for the
//   magnitudes; every bound is rounded up.
#define GM_N 38
#define GM_TR 32
#define GM_TC 64
#define GM_LD (GM_TC + GM_N - 1 + 1)
constexpr double kU32 = 5.9604644775390625e-08;  // 2^-24

__device__ __forceinline__ float bound_of(double e, double maxwin, double wsum, double coef) {
  const double d = coef * maxwin * wsum + 1e-14 * e;
  return (float)(d * 1.000001);
}

// the FP64 kernel's layout (32 x 64 outputs of one filter per block, 8
// outputs per thread), each group of 8 taps in FP32 with fused
// multiply-adds and the groups added in FP64
__global__ void __launch_bounds__(256) k_gabor_conv_f32(const Img *imgs, const float *kernels, const double *wsum,
                                                        double *const *energy, float *const *bound) {
  __shared__ float tile[GM_TR + GM_N - 1][GM_LD];
  __shared__ float2 wk[GM_N * GM_N];
  __shared__ float vmax[GM_TR][GM_TC + GM_N - 1];
  const Img im = imgs[blockIdx.z];
  const int f = blockIdx.y % 8;
  const int tilesx = (im.w + GM_TC - 1) / GM_TC, tilesy = (im.h + GM_TR - 1) / GM_TR;
  if ((int)blockIdx.x >= tilesx * tilesy) return;
  const int b0 = (blockIdx.x / tilesx) * GM_TR, a0 = (blockIdx.x % tilesx) * GM_TC;
  const int t = threadIdx.x;
  const float *K = kernels + (long)f * GM_N * GM_N * 2;
  for (int e = t; e < GM_N * GM_N; e += 256) wk[e] = make_float2(K[2 * e], K[2 * e + 1]);
  for (int r = t / 64; r < GM_TR + GM_N - 1; r += 4) {
    const int y = b0 - 18 + r;
    for (int c = t % 64; c < GM_TC + GM_N - 1; c += 64) {
      const int x = a0 - 18 + c;
      tile[r][c] = (y >= 0 && y < im.h && x >= 0 && x < im.w) ? (float)im.p[(long)y * im.w + x] : 0.0f;
    }
  }
  __syncthreads();
  // window maxima of |x|: over the 38 rows of each output row, then the 38 columns
  for (int e = t; e < GM_TR * (GM_TC + GM_N - 1); e += 256) {
    const int rr = e / (GM_TC + GM_N - 1), c = e - rr * (GM_TC + GM_N - 1);
    float m = 0.0f;
    for (int k = 0; k < GM_N; k++) m = fmaxf(m, fabsf(tile[rr + k][c]));
    vmax[rr][c] = m;
  }
  const int ty = t / 8, tx = t % 8;
  const int b = b0 + ty, a = a0 + tx * 8;
  double re[8], imv[8];
#pragma unroll
  for (int q = 0; q < 8; q++) re[q] = imv[q] = 0.0;
  const int rb = ty + 37, cb = tx * 8 + 37;
  for (int j = 0; j < GM_N; j++) {
    const float *row = &tile[rb - j][0];
    const float2 *wr = &wk[j * GM_N];
#pragma unroll
    for (int i0 = 0; i0 < GM_N; i0 += 8) {
      float r32[8], i32[8];
#pragma unroll
      for (int q = 0; q < 8; q++) r32[q] = i32[q] = 0.0f;
      const int D = (GM_N - i0 < 8) ? GM_N - i0 : 8;
      float u[15];
#pragma unroll
      for (int k = 0; k < 15; k++) u[k] = (k >= 8 - D) ? row[cb - i0 - 7 + k] : 0.0f;
#pragma unroll
      for (int di = 0; di < 8; di++) {
        if (i0 + di < GM_N) {
          const float2 w = wr[i0 + di];
#pragma unroll
          for (int q = 0; q < 8; q++) {
            const float v = u[q - di + 7];
            r32[q] = fmaf(v, w.x, r32[q]);
            i32[q] = fmaf(v, w.y, i32[q]);
          }
        }
      }
#pragma unroll
      for (int q = 0; q < 8; q++) {
        re[q] += (double)r32[q];
        imv[q] += (double)i32[q];
      }
    }
  }
  __syncthreads();  // vmax complete
  const double ws = wsum[f];
  double *E = energy[blockIdx.z] + (long)f * im.h * im.w;
  float *Bd = bound[blockIdx.z] + (long)f * im.h * im.w;
  if (b < im.h)
#pragma unroll
    for (int q = 0; q < 8; q++)
      if (a + q < im.w) {
        float m = 0.0f;
        for (int k = 0; k < GM_N; k++) m = fmaxf(m, vmax[ty][tx * 8 + q + k]);
        // |x| <= |fl32(x)| (1 + u): folded into the margin of the coefficient
        const double e = sqrt(re[q] * re[q] + imv[q] * imv[q]);
        E[(long)b * im.w + a + q] = e;
        Bd[(long)b * im.w + a + q] = bound_of(e, (double)m * (1.0 + 2 * kU32), ws, 11.0 * kU32 + 2e-13);
      }
}

// ---------------------------------------------------------------- Gabor BF16
// Implicit GEMM: M = 16 consecutive output pixels of a row, N = 16 = the 8
// filters' real and imaginary parts, K = the taps of one kernel row in
// three blocks of 16 (38 padded to 48 with zero weights).
// E(b, a) = |sum_{j,i} I(b+19-j, a+19-i) w(j,i)|.  Block = 16 rows x 32
// columns of outputs, all 8 filters; wave w owns rows 4w..4w+3 (8
// M-tiles).  The input tile holds each pixel as bf16 hi | lo in one LDS
// word, from 28 columns left of the block (taps up to 47) to 18 right;
// weights (hi, mid and lo planes, [j][i-block][n][16]) come from L2.  Per
// (kernel row, tap block) the five products go into an FP32 accumulator
// started at zero, which is then added to FP64 sums.
#define GX_TR 16
#define GX_TC 32
#define GX_W (GX_TC + 47)      // tile columns: x from a0 - 28 to a0 + 50
#define GX_LD (GX_W + 1)
#define GX_H (GX_TR + GM_N - 1)

__global__ void __launch_bounds__(256) k_gabor_conv_bf16x3(const Img *imgs, const unsigned short *w_hi,
                                                           const unsigned short *w_md, const unsigned short *w_lo,
                                                           const double *wsum, double *const *energy,
                                                           float *const *bound) {
  __shared__ unsigned tile[GX_H][GX_LD];
  __shared__ float vmax[GX_TR][GX_W];
  __shared__ unsigned eps_bits;
  const Img im = imgs[blockIdx.z];
  const int tilesx = (im.w + GX_TC - 1) / GX_TC, tilesy = (im.h + GX_TR - 1) / GX_TR;
  if ((int)blockIdx.x >= tilesx * tilesy) return;
  const int b0 = (blockIdx.x / tilesx) * GX_TR, a0 = (blockIdx.x % tilesx) * GX_TC;
  const int t = threadIdx.x;
  if (t == 0) eps_bits = 0u;
  __syncthreads();
  float eps_mine = 0.0f;  // largest |x - (xh + xl)| / |x| of this thread's pixels
  for (int e = t; e < GX_H * GX_W; e += 256) {
    const int r = e / GX_W, c = e - r * GX_W;
    const int y = b0 - 18 + r, x = a0 - 28 + c;
    const double v64 = (y >= 0 && y < im.h && x >= 0 && x < im.w) ? im.p[(long)y * im.w + x] : 0.0;
    const unsigned sp = bf16_split((float)v64);
    tile[r][c] = sp;
    const double rep = (double)__uint_as_float(sp << 16) + (double)__uint_as_float(sp & 0xffff0000u);
    if (v64 != 0.0) eps_mine = fmaxf(eps_mine, (float)(fabs(v64 - rep) / fabs(v64)) * 1.001f);
  }
  atomicMax(&eps_bits, __float_as_uint(eps_mine));  // non-negative floats order as their bits
  __syncthreads();
  // window maxima of |x| (from the split words, |xh + xl| >= |x| (1 - 2^-15))
  for (int e = t; e < GX_TR * (GX_W - 10); e += 256) {
    const int rr = e / (GX_W - 10), c = 10 + e - rr * (GX_W - 10);
    float m = 0.0f;
    for (int k = 0; k < GM_N; k++) {
      const unsigned sp = tile[rr + k][c];
      m = fmaxf(m, fabsf(__uint_as_float(sp << 16) + __uint_as_float(sp & 0xffff0000u)));
    }
    vmax[rr][c] = m;
  }
  const int wave = t >> 6, lane = t & 63;
  const int m = lane & 15, g = lane >> 4;
  double acc[8][4];
#pragma unroll
  for (int q = 0; q < 8; q++)
#pragma unroll
    for (int r = 0; r < 4; r++) acc[q][r] = 0.0;
  for (int j = 0; j < GM_N; j++) {
#pragma unroll
    for (int ib = 0; ib < 3; ib++) {
      // B: rows k = 4g..4g+3 of column n = m
      const long wo = (((long)(j * 3 + ib) * 16 + m) * 16) + 4 * g;
      const uint2 bh = *(const uint2 *)(w_hi + wo), bm = *(const uint2 *)(w_md + wo),
                  bl = *(const uint2 *)(w_lo + wo);
      bf16x4 Bh, Bm, Bl;
      Bh[0] = (short)(bh.x & 0xffff); Bh[1] = (short)(bh.x >> 16); Bh[2] = (short)(bh.y & 0xffff); Bh[3] = (short)(bh.y >> 16);
      Bm[0] = (short)(bm.x & 0xffff); Bm[1] = (short)(bm.x >> 16); Bm[2] = (short)(bm.y & 0xffff); Bm[3] = (short)(bm.y >> 16);
      Bl[0] = (short)(bl.x & 0xffff); Bl[1] = (short)(bl.x >> 16); Bl[2] = (short)(bl.y & 0xffff); Bl[3] = (short)(bl.y >> 16);
      f32x4 part[8];
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const int rr = wave * 4 + (q >> 1), seg = q & 1;
        // A: row m of the M-tile, taps i = 16 ib + 4g + u: tile column 16 seg + m + 47 - i
        const unsigned *src = &tile[rr + 37 - j][16 * seg + m + 47 - 16 * ib - 4 * g];
        bf16x4 Ah, Al;
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const unsigned v = src[-u];
          Ah[u] = (short)(v & 0xffff);
          Al[u] = (short)(v >> 16);
        }
        f32x4 z = {0.f, 0.f, 0.f, 0.f};
        z = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Ah, Bl, z, 0, 0, 0);
        z = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Al, Bm, z, 0, 0, 0);
        z = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Ah, Bm, z, 0, 0, 0);
        z = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Al, Bh, z, 0, 0, 0);
        z = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Ah, Bh, z, 0, 0, 0);
        part[q] = z;
      }
#pragma unroll
      for (int q = 0; q < 8; q++)
#pragma unroll
        for (int r = 0; r < 4; r++) acc[q][r] += (double)part[q][r];
    }
  }
  __syncthreads();  // vmax and eps complete
  const double coef = 85.0 * kU32 + (double)__uint_as_float(eps_bits) + 2e-13;
  // D[i][n]: lane holds n = m, pixels i = 4g..4g+3; the partner lane m ^ 1
  // holds the other part of the same filter
  double *E0 = energy[blockIdx.z];
  float *B0 = bound[blockIdx.z];
  const long plane = (long)im.h * im.w;
  const int f = m >> 1;
  const double ws = wsum[f];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int rr = wave * 4 + (q >> 1);
    const int b = b0 + rr;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const double mine = acc[q][r];
      const double other = __shfl_xor(mine, 1);
      const int cc = 16 * (q & 1) + 4 * g + r, a = a0 + cc;
      if (!(m & 1) && b < im.h && a < im.w) {
        float mw = 0.0f;
        for (int k = 0; k < GM_N; k++) mw = fmaxf(mw, vmax[rr][cc + 10 + k]);
        const double e = sqrt(mine * mine + other * other);
        E0[f * plane + (long)b * im.w + a] = e;
        B0[f * plane + (long)b * im.w + a] = bound_of(e, (double)mw * (1.0 + 6.2e-5), ws, coef);
      }
    }
  }
}

void launch_gabor_mixed(int precision, const Img *imgs, const void *k0, const void *k1, const void *k2,
                        const double *wsum, const double *taps64, double *const *energy, float *const *bound,
                        double *params, int *flags, OutRef o, int nimg, int maxh, int maxw, hipStream_t s) {
  if (precision == 1) {
    const int tiles = ((maxw + GM_TC - 1) / GM_TC) * ((maxh + GM_TR - 1) / GM_TR);
    k_gabor_conv_f32<<<dim3(tiles, 8, nimg), 256, 0, s>>>(imgs, (const float *)k0, wsum, energy, bound);
  } else {
    const int tiles = ((maxw + GX_TC - 1) / GX_TC) * ((maxh + GX_TR - 1) / GX_TR);
    k_gabor_conv_bf16x3<<<dim3(tiles, 1, nimg), 256, 0, s>>>(imgs, (const unsigned short *)k0,
                                                             (const unsigned short *)k1, (const unsigned short *)k2,
                                                             wsum, energy, bound);
  }
  launch_gabor_post_exact(imgs, energy, bound, taps64, params, flags, false, o, nimg, maxh, maxw, s);
}

}  // namespace wnd
