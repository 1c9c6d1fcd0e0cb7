// Mixed-radix Stockham FFT of one line held in LDS (shared by the feature
// path and the adjoint kernels).  Every radix is evaluated as a direct
// R-point DFT with combined twiddles from a length-n table
// tw[k] = exp(-2 pi i k / n).  All threads of the block take part; on
// return *res points at the buffer (a or b) that holds the result.
#pragma once
#include "kernels.hpp"

namespace wnd {

// One line per group of `tpl` threads (lane = index in the group); every
// group of the block runs the same plan, so the barriers are uniform.
__device__ inline void fft_lds_group(double2 *a, double2 *b, const FftPlan &pl, int lane, int tpl, double2 **res) {
  const int n = pl.n;
  int Ns = 1;
  double2 *src = a, *dst = b;
  for (int s = 0; s < pl.nst; s++) {
    const int R = pl.radix[s];
    const int nr = n / R, step = n / (Ns * R);
    for (int t = lane; t < n; t += tpl) {
      const int j = t % nr, q = t / nr;
      const int k = j % Ns;
      const int base = k * step + q * nr;
      double re = 0, im = 0;
      long idx = 0;
      for (int r = 0; r < R; r++) {
        const double2 v = src[j + r * nr];
        const double2 w = pl.tw[idx];
        re += v.x * w.x - v.y * w.y;
        im += v.x * w.y + v.y * w.x;
        idx += base;
        if (idx >= n) idx -= n;
      }
      dst[(j / Ns) * Ns * R + k + q * Ns] = make_double2(re, im);
    }
    __syncthreads();
    double2 *tt = src; src = dst; dst = tt;
    Ns *= R;
  }
  *res = src;
}

__device__ inline void fft_lds(double2 *a, double2 *b, const FftPlan &pl, double2 **res) {
  fft_lds_group(a, b, pl, threadIdx.x, blockDim.x, res);
}

// lines per 256-thread block for a line length n (lines * n <= 1024 up to n = 1024)
__host__ __device__ inline int fft_lines(int n) { return n <= 256 ? 4 : (n <= 512 ? 2 : 1); }

}  // namespace wnd
