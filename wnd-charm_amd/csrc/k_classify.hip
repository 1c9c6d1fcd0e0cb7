// WND classifier kernels: min/max normalisation, Fisher scores with the
// top-fraction threshold, WND5 marginal probabilities.  FP64; per-feature
// and per-sample loops run in the reference's order (TrainingSet.cpp).
#include "common.hpp"

namespace wnd {

// TrainingSet::normalize (TrainingSet.cpp:2090-2123): one thread per feature
__global__ void k_norm_train(double *data, int n, int f, double *mins, double *maxs) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= f) return;
  double mn = kFltMax, mx = -kFltMax;
  for (int i = 0; i < n; i++) {
    const double v = data[(long)i * f + s];
    if (!moments_ok(v)) continue;
    if (v > mx) mx = v;
    if (v < mn) mn = v;
  }
  maxs[s] = mx;
  mins[s] = mn;
  for (int i = 0; i < n; i++) {
    double v = data[(long)i * f + s];
    if (v != v || v < mn || (mx - mn) < kDblEps) v = 0;
    else if (v > mx) v = 100;
    else v = 100 * ((v - mn) / (mx - mn));
    data[(long)i * f + s] = v;
  }
}

// signatures::normalize (signatures.cpp:196-215)
__global__ void k_norm_test(double *vec, int nvec, int f, const double *mins, const double *maxs) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)nvec * f) return;
  const int s = e % f;
  double v = vec[e];
  const double mn = mins[s], mx = maxs[s];
  if (v != v || v < mn || (mx - mn) < kDblEps) v = 0;
  else if (v > mx) v = 100;
  else v = 100 * ((v - mn) / (mx - mn));
  vec[e] = v;
}

// SetFisherScores (TrainingSet.cpp:2183-2250): per-class Moments2 in sample order
__global__ void k_fisher(const double *data, const int *cls, int n, int f, int C, double *w) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= f) return;
  double gm = 0;
  size_t gn = 0;
  double dev = 0, var = 0;
  // the feature-wide mean (Moments2 over class != 0 samples)
  for (int i = 0; i < n; i++) {
    if (!cls[i]) continue;
    const double v = data[(long)i * f + s];
    if (!moments_ok(v)) continue;
    const size_t n1 = gn;
    (void)n1;
    gn = gn + 1;
    const double delta = v - gm;
    const double delta_n = delta / gn;
    gm = gm + delta_n;
  }
  for (int c = 1; c <= C; c++) {
    double mean = 0, M2 = 0;
    size_t cn = 0;
    for (int i = 0; i < n; i++) {
      if (cls[i] != c) continue;
      const double v = data[(long)i * f + s];
      if (!moments_ok(v)) continue;
      const size_t n1 = cn;
      cn = cn + 1;
      const double delta = v - mean;
      const double delta_n = delta / cn;
      const double term1 = delta * delta_n * n1;
      mean = mean + delta_n;
      M2 += term1;
    }
    const double d = mean - gm;
    dev += d * d;
    var += cn > 2 ? (M2 / (cn - 1)) : 0.0;
  }
  if (C > 1) {
    dev /= (C - 1);
    var /= C;
  } else dev = 0;
  if (var < 0.000001) var = 0.000001;
  w[s] = dev / var;
}

// threshold = value at descending rank floor(used*f + 0.5); weights below are zeroed
__global__ void __launch_bounds__(1024) k_fisher_threshold(double *w, int f, int last) {
  __shared__ double thr;
  if (threadIdx.x == 0) thr = -1.79769313486231570815e+308;
  __syncthreads();
  if (last < f) {
    for (int i = threadIdx.x; i < f; i += blockDim.x) {
      const double wi = w[i];
      int rank = 0;
      for (int j = 0; j < f; j++) {
        const double wj = w[j];
        rank += (wj > wi) || (wj == wi && j < i);
      }
      if (rank == last) thr = wi;
    }
  }
  __syncthreads();
  const double t = thr;
  for (int i = threadIdx.x; i < f; i += blockDim.x)
    if (w[i] < t) w[i] = 0.0;
}

// classify2 distances (TrainingSet.cpp:2445-2495): one thread per (test, train)
__global__ void k_wnd_dist(const double *train, const double *wts, const double *test, int n, int f, int ntest,
                           double *dist) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)n * ntest) return;
  const int q = e / n, i = e - (long)q * n;
  const double *tv = test + (long)q * f, *sv = train + (long)i * f;
  double ds = 0.0;
  for (int s = 0; s < f; s++) {
    const double ws = wts[s];
    if (ws < kDblEps) continue;
    const double d = fabs(tv[s] - sv[s]);
    if (d < kDblEps) continue;
    ds += (ws * ws) * (d * d);
  }
  dist[e] = ds;
}

// per test vector: class means of dist^-5 in sample order, probabilities, argmax
__global__ void k_wnd_classes(const double *dist, const int *cls, int n, int C, int ntest, double *prob,
                              double *norm, long *pred) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ntest) return;
  double csim[64];
  int nper[64];
  for (int c = 0; c <= C; c++) { csim[c] = 0; nper[c] = 0; }
  for (int i = 0; i < n; i++) {
    const double ds = dist[(long)q * n + i];
    const int ci = cls[i];
    if (ds < kDblEps || ci < 0 || ci > C) continue;  // labels are validated on the host; never index outside
    csim[ci] += pow(ds, -5.0);
    nper[ci]++;
  }
  long best = -1;
  double maxs = 0;
  for (int c = 1; c <= C; c++) {
    if (nper[c] == 0) continue;
    csim[c] /= nper[c];
    if (csim[c] > maxs) { maxs = csim[c]; best = c; }
  }
  double sum = 0;
  for (int c = 1; c <= C; c++) sum += csim[c];
  for (int c = 1; c <= C; c++) prob[(long)q * C + c - 1] = csim[c] / sum;
  norm[q] = sum;
  pred[q] = best;
}

void launch_norm_train(double *data, int n, int f, double *mins, double *maxs, hipStream_t s) {
  k_norm_train<<<(f + 127) / 128, 128, 0, s>>>(data, n, f, mins, maxs);
}
void launch_norm_test(double *vec, int nvec, int f, const double *mins, const double *maxs, hipStream_t s) {
  const long n = (long)nvec * f;
  k_norm_test<<<(n + 255) / 256, 256, 0, s>>>(vec, nvec, f, mins, maxs);
}
void launch_fisher(const double *data, const int *cls, int n, int f, int C, double used, double *w, hipStream_t s) {
  k_fisher<<<(f + 127) / 128, 128, 0, s>>>(data, cls, n, f, C, w);
  const int last = (int)floor((used * (double)f) + 0.5);
  k_fisher_threshold<<<1, 1024, 0, s>>>(w, f, last);
}
void launch_wnd5(const double *train, const int *cls, const double *w, const double *test, int n, int f, int C,
                 int ntest, double *dist, double *prob, double *norm, long *pred, hipStream_t s) {
  const long e = (long)n * ntest;
  k_wnd_dist<<<(e + 255) / 256, 256, 0, s>>>(train, w, test, n, f, ntest, dist);
  k_wnd_classes<<<(ntest + 63) / 64, 64, 0, s>>>(dist, cls, n, C, ntest, prob, norm, pred);
}

}  // namespace wnd
