// Work-group LSD radix sort of 64-bit keys (1024 threads, 8-bit digits),
// stable ranking inside 1024-key tiles through wave ballots.  Digit passes
// in which every key has the same digit are skipped.
#pragma once
#include "common.hpp"

namespace wnd {

// Sorts keys k0[0..n) using k1[0..n) as the ping-pong buffer; returns the
// buffer that holds the sorted keys.  Requires blockDim.x == 1024.  The
// eight digit histograms are counted in one read of the keys; each 4096-key
// tile's keys are loaded while the previous tile is ranked.
__device__ uint64_t *block_radix_sort(uint64_t *k0p, uint64_t *k1p, long n) {
  gptr<uint64_t> k0 = to_global(k0p), k1 = to_global(k1p);  // global, not FLAT: LDS waits leave them in flight
  __shared__ int hist[8][256];
  __shared__ int base[256];
  __shared__ int tot[256];
  __shared__ int wcnt[16][256];
  __shared__ int skip;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int e = threadIdx.x; e < 8 * 256; e += 1024) (&hist[0][0])[e] = 0;
  __syncthreads();
  for (long i = threadIdx.x; i < n; i += 1024) {
    const uint64_t key = k0[i];
#pragma unroll
    for (int p = 0; p < 8; p++) atomicAdd(&hist[p][(key >> (8 * p)) & 255], 1);
  }
  __syncthreads();
  for (int pass = 0; pass < 8; pass++) {
    const int sh = pass * 8;
    __syncthreads();  // every thread has read the previous pass's skip flag
    if (threadIdx.x == 0) {
      int acc = 0, sk = 0;
      for (int d = 0; d < 256; d++) {
        base[d] = acc;
        acc += hist[pass][d];
        if (hist[pass][d] == n) sk = 1;
      }
      skip = sk || n <= 1;
    }
    __syncthreads();
    if (skip) continue;
    // tiles of 4096 keys; wave w ranks keys t0 + 256w .. +255 in four
    // ballot rounds of 64 (stable: round, then lane), counting per digit in
    // its own row of wcnt
    uint64_t next[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const long i = 256 * wv + 64 * r + lane;
      next[r] = i < n ? k0[i] : 0;
    }
    for (long t0 = 0; t0 < n; t0 += 4096) {
      uint64_t key[4];
      int dg[4], lrank[4];
#pragma unroll
      for (int r = 0; r < 4; r++) {
        key[r] = next[r];
        const long i = t0 + 4096 + 256 * wv + 64 * r + lane;
        if (i < n) next[r] = k0[i];  // the next tile's keys, in flight meanwhile
      }
      for (int d = lane; d < 256; d += 64) wcnt[wv][d] = 0;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const long i = t0 + 256 * wv + 64 * r + lane;
        const bool valid = i < n;
        dg[r] = valid ? (int)((key[r] >> sh) & 255) : 256;
        uint64_t peers = __ballot(1);
        for (int b = 0; b < 9; b++) {
          const uint64_t bb = __ballot((dg[r] >> b) & 1);
          peers &= ((dg[r] >> b) & 1) ? bb : ~bb;
        }
        const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
        const int rank = __popcll(peers & lt);
        const int pos = valid ? wcnt[wv][dg[r]] : 0;
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (valid && rank == 0) wcnt[wv][dg[r]] = pos + __popcll(peers);
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        lrank[r] = pos + rank;
      }
      __syncthreads();
      if (threadIdx.x < 256) {
        int acc = 0;
        for (int w = 0; w < 16; w++) {
          const int c = wcnt[w][threadIdx.x];
          wcnt[w][threadIdx.x] = acc;
          acc += c;
        }
        tot[threadIdx.x] = acc;
      }
      __syncthreads();
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (dg[r] < 256) k1[base[dg[r]] + wcnt[wv][dg[r]] + lrank[r]] = key[r];
      __syncthreads();
      if (threadIdx.x < 256) base[threadIdx.x] += tot[threadIdx.x];
    }
    __syncthreads();
    gptr<uint64_t> t = k0;
    k0 = k1;
    k1 = t;
    uint64_t *tp = k0p;
    k0p = k1p;
    k1p = tp;
  }
  __syncthreads();
  return k0p;
}

}  // namespace wnd
