// Work-group LSD radix sort of 64-bit keys (1024 threads, 8-bit digits),
// stable ranking inside 1024-key tiles through wave ballots.  Digit passes
// in which every key has the same digit are skipped.
#pragma once
#include "common.hpp"

namespace wnd {

// Sorts keys k0[0..n) using k1[0..n) as the ping-pong buffer; returns the
// buffer that holds the sorted keys.  Requires blockDim.x == 1024.
__device__ uint64_t *block_radix_sort(uint64_t *k0, uint64_t *k1, long n) {
  __shared__ int hist[256];
  __shared__ int base[256];
  __shared__ int wcnt[16][256];
  __shared__ int skip;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int pass = 0; pass < 8; pass++) {
    const int sh = pass * 8;
    if (threadIdx.x < 256) hist[threadIdx.x] = 0;
    __syncthreads();
    for (long i = threadIdx.x; i < n; i += 1024) atomicAdd(&hist[(k0[i] >> sh) & 255], 1);
    __syncthreads();
    if (threadIdx.x == 0) {
      int acc = 0, sk = 0;
      for (int d = 0; d < 256; d++) {
        base[d] = acc;
        acc += hist[d];
        if (hist[d] == n) sk = 1;
      }
      skip = sk || n <= 1;
    }
    __syncthreads();
    if (skip) continue;
    for (long t0 = 0; t0 < n; t0 += 1024) {
      for (int e = threadIdx.x; e < 16 * 256; e += 1024) (&wcnt[0][0])[e] = 0;
      __syncthreads();
      const long i = t0 + threadIdx.x;
      const bool valid = i < n;
      const uint64_t key = valid ? k0[i] : 0;
      const int dg = valid ? (int)((key >> sh) & 255) : 256;
      uint64_t peers = __ballot(1);
      for (int b = 0; b < 9; b++) {
        const uint64_t bb = __ballot((dg >> b) & 1);
        peers &= ((dg >> b) & 1) ? bb : ~bb;
      }
      const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
      const int rank = __popcll(peers & lt);
      if (valid && rank == 0) wcnt[wv][dg] = __popcll(peers);
      __syncthreads();
      if (threadIdx.x < 256) {
        int acc = 0;
        for (int w = 0; w < 16; w++) {
          const int c = wcnt[w][threadIdx.x];
          wcnt[w][threadIdx.x] = acc;
          acc += c;
        }
        hist[threadIdx.x] = acc;
      }
      __syncthreads();
      if (valid) k1[base[dg] + wcnt[wv][dg] + rank] = key;
      __syncthreads();
      if (threadIdx.x < 256) base[threadIdx.x] += hist[threadIdx.x];
      __syncthreads();
    }
    uint64_t *t = k0;
    k0 = k1;
    k1 = t;
  }
  __syncthreads();
  return k0;
}

}  // namespace wnd
