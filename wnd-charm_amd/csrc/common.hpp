// Shared device-side types and helpers for the MI355X feature engine.
//
// Every kernel is batch-generic: it receives a device array of image
// descriptors (one per image of the batch, sizes may differ) and a
// grid whose y dimension indexes the image.  Blocks beyond an image's own
// extent exit immediately, so one launch serves a size-bucketed batch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wnd {

// Global-memory view of a device pointer.  Pointers read from descriptor
// structs are generic to the compiler, so their accesses become FLAT
// instructions, which count against the LDS wait counter as well: a kernel
// that mixes LDS and global traffic would wait for its global loads at
// every LDS access.  Kernels address global memory through gptr instead.
template <class T>
using gptr = __attribute__((address_space(1))) T *;
template <class T>
__device__ __forceinline__ gptr<T> to_global(T *p) {
  return (gptr<T>)p;
}

// One FP64 row-major plane on the device (an ImageMatrix pixel plane).
struct Img {
  double *p;
  int h, w;
};

// Pixel statistics of a plane (Moments2 semantics: NaN and |x| > FLT_MAX
// values are not counted; var/std are 0 for n <= 2).
struct Stats {
  double mn, mx, mean, var, stdv;
  long long n;
};

// Output addressing: rows[i] is where entry i of a launch writes its
// feature group (result row of its image + the group's column offset), so
// one launch can serve a feature group on several DAG nodes at once.
struct OutRef {
  double *const *rows;
};

// XCD-aware block mapping.  MI355X deals consecutive workgroups round-robin
// to its 8 XCDs, each with its own L2.  For a 1-D grid of 8 * per_item *
// ceil(items / 8) blocks, this gives every block of one item the same XCD
// (item = its blocks' shared input, e.g. one image plane), so the item is
// fetched into one L2 only.  Returns false for padding blocks.
__device__ __forceinline__ bool xcd_item(int per_item, int items, int &item, int &sub) {
  const int b = blockIdx.x, xcd = b & 7, k = b >> 3;
  sub = k % per_item;
  item = (k / per_item) * 8 + xcd;
  return item < items;
}
inline int xcd_grid(int per_item, int items) { return 8 * per_item * ((items + 7) / 8); }

constexpr double kFltMax = 3.40282346638528859812e+38;
constexpr double kDblEps = 2.2204460492503131e-16;
constexpr double kInf = 10E200;  // the reference's INF (src/cmatrix.h:48)

__device__ __forceinline__ bool moments_ok(double x) {
  return !(x != x) && !(x > kFltMax) && !(x < -kFltMax);
}

// histogram bin exactly as ImageMatrix::histogram (src/cmatrix.cpp:1069-1103):
// bin = (unsigned long)((v - min) * scale), clamped to nbins-1.
__device__ __forceinline__ int hist_bin(double v, double mn, double scale, int nbins) {
  double t = (v - mn) * scale;
  // (unsigned long) of a negative or NaN value is undefined in C; the
  // reference only ever feeds v >= min.  Values >= nbins clamp.
  unsigned long long b = (t >= 18446744073709551615.0 || t != t) ? (unsigned long long)nbins
                                                                 : (unsigned long long)t;
  if (b >= (unsigned long long)nbins) b = nbins - 1;
  return (int)b;
}

__device__ __forceinline__ double hist_scale(double mn, double mx, int nbins) {
  return (mx - mn > 0) ? (double)nbins / (double)(mx - mn) : 0.0;
}

// Order-preserving map of a double to uint64 (for radix sort / select).
__device__ __forceinline__ uint64_t dkey(double v) {
  uint64_t u = __double_as_longlong(v);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double dkey_inv(uint64_t k) {
  uint64_t u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double(u);
}

// Block-wide reductions (blockDim.x a multiple of 64, <= 1024).
template <typename T, typename Op>
__device__ T block_reduce(T v, Op op, T *sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_down(v, o, 64));
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  if (wid == 0) {
    v = (lane < nw) ? sh[lane] : sh[0];
    if (lane >= nw) v = sh[0];
    for (int o = 32; o > 0; o >>= 1) {
      T u = __shfl_down(v, o, 64);
      if (lane + o < nw) v = op(v, u);
    }
    if (lane == 0) sh[0] = v;
  }
  __syncthreads();
  T r = sh[0];
  __syncthreads();
  return r;
}

struct SumOp { template <typename T> __device__ T operator()(T a, T b) const { return a + b; } };
struct MinOp { template <typename T> __device__ T operator()(T a, T b) const { return a < b ? a : b; } };
struct MaxOp { template <typename T> __device__ T operator()(T a, T b) const { return a > b ? a : b; } };

}  // namespace wnd
