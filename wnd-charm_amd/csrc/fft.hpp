// Mixed-radix Stockham FFT of one line held in LDS (shared by the feature
// path and the adjoint kernels).  Every radix is evaluated as a direct
// R-point DFT with combined twiddles from a length-n table
// tw[k] = exp(-2 pi i k / n).  All threads of the block take part; on
// return *res points at the buffer (a or b) that holds the result.
#pragma once
#include "kernels.hpp"

namespace wnd {

// Exact t / d for t < 2^20, d <= 4096 as a multiply-high: with
// M = floor((2^32 - 1) / d) + 1 (2^32 / d for powers of two),
// floor(t M / 2^32) = floor(t / d) because t (M - 2^32/d) < 2^32 / d.
__device__ __forceinline__ unsigned long long fft_magic(unsigned d) { return (unsigned long long)(0xFFFFFFFFu / d) + 1; }
__device__ __forceinline__ unsigned fft_div(unsigned t, unsigned long long m) {
  return (unsigned)(((unsigned long long)t * m) >> 32);
}

// One radix-R stage for R in {2, 3, 4, 5, 7}: one thread per butterfly j
// loads its R inputs once and forms the R outputs q; each output is the same
// sum, in the same order and with the same twiddles, as in the per-output
// form below.
template <int R, class TW>
__device__ __forceinline__ void fft_stage(const double2 *src, double2 *dst, TW tw, int n, int Ns, int lane,
                                          int tpl) {
  const int nr = n / R, step = n / (Ns * R);
  const unsigned long long mns = fft_magic(Ns);
  for (int j = lane; j < nr; j += tpl) {
    const int jq = fft_div(j, mns), k = j - jq * Ns;
    double2 v[R];
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = src[j + r * nr];
#pragma unroll
    for (int q = 0; q < R; q++) {
      const int base = k * step + q * nr;
      double re = 0, im = 0;
      int idx = 0;
#pragma unroll
      for (int r = 0; r < R; r++) {
        const double2 w = tw[idx];
        re += v[r].x * w.x - v[r].y * w.y;
        im += v[r].x * w.y + v[r].y * w.x;
        idx += base;
        if (idx >= n) idx -= n;
      }
      dst[jq * Ns * R + k + q * Ns] = make_double2(re, im);
    }
  }
}

// One line per group of `tpl` threads (lane = index in the group); every
// group of the block runs the same plan, so the barriers are uniform.  tw:
// the plan's twiddles, in global memory or staged in LDS by the caller.
template <class TW>
__device__ inline void fft_lds_group_tw(double2 *a, double2 *b, const FftPlan &pl, TW tw, int lane, int tpl,
                                        double2 **res) {
  const int n = pl.n;
  int Ns = 1;
  double2 *src = a, *dst = b;
  for (int s = 0; s < pl.nst; s++) {
    const int R = pl.radix[s];
    switch (R) {
      case 2: fft_stage<2>(src, dst, tw, n, Ns, lane, tpl); break;
      case 3: fft_stage<3>(src, dst, tw, n, Ns, lane, tpl); break;
      case 4: fft_stage<4>(src, dst, tw, n, Ns, lane, tpl); break;
      case 5: fft_stage<5>(src, dst, tw, n, Ns, lane, tpl); break;
      case 7: fft_stage<7>(src, dst, tw, n, Ns, lane, tpl); break;
      default: {  // other primes: one output per thread, as a direct R-point DFT
        const int nr = n / R, step = n / (Ns * R);
        const unsigned long long mnr = fft_magic(nr), mns = fft_magic(Ns);
        for (int t = lane; t < n; t += tpl) {
          const int q = fft_div(t, mnr), j = t - q * nr;
          const int jq = fft_div(j, mns), k = j - jq * Ns;
          const int base = k * step + q * nr;
          double re = 0, im = 0;
          int idx = 0;
          for (int r = 0; r < R; r++) {
            const double2 v = src[j + r * nr];
            const double2 w = tw[idx];
            re += v.x * w.x - v.y * w.y;
            im += v.x * w.y + v.y * w.x;
            idx += base;
            if (idx >= n) idx -= n;
          }
          dst[jq * Ns * R + k + q * Ns] = make_double2(re, im);
        }
      }
    }
    __syncthreads();
    double2 *tt = src; src = dst; dst = tt;
    Ns *= R;
  }
  *res = src;
}

__device__ inline void fft_lds_group(double2 *a, double2 *b, const FftPlan &pl, int lane, int tpl, double2 **res) {
  fft_lds_group_tw(a, b, pl, pl.tw, lane, tpl, res);
}

__device__ inline void fft_lds(double2 *a, double2 *b, const FftPlan &pl, double2 **res) {
  fft_lds_group(a, b, pl, threadIdx.x, blockDim.x, res);
}

// lines per 256-thread block for a line length n (lines * n <= 1024 up to n = 1024)
__host__ __device__ inline int fft_lines(int n) { return n <= 256 ? 4 : (n <= 512 ? 2 : 1); }

}  // namespace wnd
