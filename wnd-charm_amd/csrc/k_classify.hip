// WND classifier kernels: min/max normalisation, Fisher scores with the
// top-fraction threshold, WND5 marginal probabilities.  FP64; per-feature
// and per-sample loops run in the reference's order (TrainingSet.cpp).
#include <algorithm>
#include "common.hpp"
#include "kernels.hpp"

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

// WNNclassify (TrainingSet.cpp:2375-2417) with TrainingSet::distance(.., 2)
// (TrainingSet.cpp:2357-2363): one thread per (test, train) pair,
// sqrt(sum_s w_s (x_s - y_s)^2) in feature order, an identical sample
// (distance below 1/INF) counted as INF
__global__ void k_wnn_dist(const double *train, const double *wts, const double *test, int n, int f, int ntest,
                           double *dist) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)n * ntest) return;
  const int q = e / n, i = e - (long)q * n;
  const double *tv = test + (long)q * f, *sv = train + (long)i * f;
  double acc = 0.0;
  for (int s = 0; s < f; s++) {
    const double d = tv[s] - sv[s];
    acc = acc + wts[s] * (d * d);
  }
  double d = sqrt(acc);
  if (d < 1 / kInf) d = kInf;
  dist[e] = d;
}

// per test vector, samples in order: the first strictly closest sample's
// class, each class's closest distance, marginals 1/d over their sum
__global__ void k_wnn_classes(const double *dist, const int *cls, int n, int C, int ntest, double *prob,
                              double *norm, long *pred) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ntest) return;
  double p[64];
  for (int c = 0; c <= C; c++) p[c] = kInf;
  long best = 0;
  double closest = kInf;
  for (int i = 0; i < n; i++) {
    const double d = dist[(long)q * n + i];
    const int ci = cls[i];
    if (ci < 0 || ci > C) continue;  // labels are validated on the host; never index outside
    if (d < closest) {
      closest = d;
      best = ci;
    }
    if (d < p[ci]) p[ci] = d;
  }
  double sum = 0;
  for (int c = 1; c <= C; c++)
    if (p[c] != 0) sum += 1 / p[c];
  for (int c = 1; c <= C; c++)
    prob[(long)q * C + c - 1] = sum == 0 ? 0.0 : (p[c] == 0 ? 1.0 : (1 / p[c]) / sum);
  norm[q] = sum;
  pred[q] = best;
}

void launch_wnn(const double *train, const int *cls, const double *w, const double *test, int n, int f, int C,
                int ntest, double *dist, double *prob, double *norm, long *pred, hipStream_t s) {
  const long e = (long)n * ntest;
  k_wnn_dist<<<(e + 255) / 256, 256, 0, s>>>(train, w, test, n, f, ntest, dist);
  k_wnn_classes<<<(ntest + 63) / 64, 64, 0, s>>>(dist, cls, n, C, ntest, prob, norm, pred);
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

// ---- the wndcharm Python package's classifier maths --------------------
// normalize_by_columns / ReplaceNonReal (wndcharm/utils.py:124-240),
// FisherFeatureWeights.NewFromFeatureSpace (wndcharm/FeatureWeights.py:
// 165-236) and FeatureSpaceClassification.NewWND5
// (wndcharm/FeatureSpacePrediction.py:641-799).  numpy's axis-0 reductions
// of a 2-D array are running sums in row order; its 1-D sums (the class
// means of the similarities) are pairwise (np_sum below); scipy's cdist
// 'sqeuclidean' is a running sum in feature order.

// np.maximum / np.minimum: NaN propagates
__device__ __forceinline__ double np_max2(double a, double b) { return a != a ? a : (b != b ? b : (a >= b ? a : b)); }
__device__ __forceinline__ double np_min2(double a, double b) { return a != a ? a : (b != b ? b : (a <= b ? a : b)); }

// flags of k_py_norm (the C ABI's WNDGPU_PY_*)
constexpr int PY_ZSCORE = 1, PY_NONREAL = 2, PY_GMINS = 4, PY_GMAXS = 8, PY_GMEANS = 16, PY_GSTDEVS = 32,
              PY_REPLACE_ONLY = 64;

// any NaN / +-inf in the matrix (ReplaceNonReal's np.any(mask)); *flag starts at 0
__global__ void k_py_any_invalid(const double *M, long total, int *flag) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  bool bad = false;
  for (long k = e; k < total; k += (long)gridDim.x * blockDim.x) bad |= !isfinite(M[k]);
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

// one thread per feature column s
__global__ void k_py_norm(double *M, int n, int f, int flags, double *mins, double *maxs, double *means,
                          double *stdevs, const int *flag) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= f) return;
  const bool any = *flag != 0;
  const bool gmn = flags & PY_GMINS, gmx = flags & PY_GMAXS;
  auto col = [&](long i) -> double & { return M[i * f + s]; };
  auto clip = [&](bool hl, double lo, bool hh, double hi) {
    for (long i = 0; i < n; i++) {
      double v = col(i);
      if (hl) v = np_max2(v, lo);
      if (hh) v = np_min2(v, hi);
      col(i) = v;
    }
  };
  // extremum over the finite entries; a column with none reads as numpy's
  // masked-array fill value 1e20
  auto masked_ext = [&](bool want_max) {
    bool seen = false;
    double r = 0;
    for (long i = 0; i < n; i++) {
      const double v = col(i);
      if (!isfinite(v)) continue;
      if (!seen || (want_max ? v > r : v < r)) r = v;
      seen = true;
    }
    return seen ? r : 1e20;
  };
  // ReplaceNonReal: its own bounds are rebound (no mins: maxs := finite max,
  // mins stays absent; mins but no maxs: mins := finite min), then clip, NaN -> 0
  auto replace = [&]() {
    bool hl = gmn, hh = gmx;
    double lo = gmn ? mins[s] : 0.0, hi = gmx ? maxs[s] : 0.0;
    if (!gmn) { hi = masked_ext(true); hh = true; }
    if (!hh) { lo = masked_ext(false); hl = true; }
    clip(hl, lo, hh, hi);
    for (long i = 0; i < n; i++)
      if (col(i) != col(i)) col(i) = 0.0;
  };
  if (flags & PY_REPLACE_ONLY) {
    if (any) replace();
    return;
  }
  if (flags & PY_NONREAL) {
    if (any) replace();
    else if (gmn || gmx) clip(gmn, gmn ? mins[s] : 0.0, gmx, gmx ? maxs[s] : 0.0);
  } else if (gmn || gmx) {
    clip(gmn, gmn ? mins[s] : 0.0, gmx, gmx ? maxs[s] : 0.0);
  }
  if (!(flags & PY_ZSCORE)) {
    double mn, mx;
    if (gmn) mn = mins[s];
    else {
      mn = col(0);
      for (long i = 1; i < n; i++) mn = np_min2(mn, col(i));
      mins[s] = mn;
    }
    if (gmx) mx = maxs[s];
    else {
      mx = col(0);
      for (long i = 1; i < n; i++) mx = np_max2(mx, col(i));
      maxs[s] = mx;
    }
    const double rng = mx - mn;
    for (long i = 0; i < n; i++) {
      double v = col(i);
      v = v - mn;   // feature_matrix -= mins
      v = v / rng;  // feature_matrix /= (maxs - mins)
      v = v * 100;  // feature_matrix *= 100
      col(i) = rng == 0 ? 0.0 : v;
    }
  } else {
    double sum = 0.0;
    for (long i = 0; i < n; i++) sum = sum + col(i);
    const double mean = sum / n;
    if (!(flags & PY_GMEANS)) means[s] = mean;
    if (!(flags & PY_GSTDEVS)) {
      double ss = 0.0;
      for (long i = 0; i < n; i++) {
        const double d = col(i) - mean;
        ss = ss + d * d;
      }
      stdevs[s] = sqrt(ss / n);
    }
    const double mu = means[s], sd = stdevs[s];
    for (long i = 0; i < n; i++) {
      double v = col(i);
      v = v - mu;
      v = v / sd;
      col(i) = sd == 0 ? 0.0 : v;
    }
  }
}

// FisherFeatureWeights.NewFromFeatureSpace, one thread per feature: the
// population mean over every row, per class (rows in order) the mean and
// the ddof=1 variance, then w = (sum_c (pop - mean_c)^2 / (C-1)) / mean_c(var_c)
// with a zero denominator read as NaN and NaN / +-inf weights set to 0
__global__ void k_py_fisher(const double *data, const int *cls, int n, int f, int C, double *w) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= f) return;
  double pop = 0.0;
  for (long i = 0; i < n; i++) pop = pop + data[i * f + s];
  pop = pop / n;
  double den = 0.0, num = 0.0;
  for (int c = 1; c <= C; c++) {
    double sum = 0.0;
    long nc = 0;
    for (long i = 0; i < n; i++)
      if (cls[i] == c) { sum = sum + data[i * f + s]; nc++; }
    const double mean = sum / (double)nc;
    double ss = 0.0;
    for (long i = 0; i < n; i++)
      if (cls[i] == c) {
        const double d = data[i * f + s] - mean;
        ss = ss + d * d;
      }
    den = den + ss / (double)(nc - 1 > 0 ? nc - 1 : 0);
    const double d = pop - mean;
    num = num + d * d;
  }
  den = den / C;
  if (den == 0) den = __longlong_as_double(0x7ff8000000000000LL);  // NaN
  const double v = (num / (double)(C - 1)) / den;
  w[s] = isfinite(v) ? v : 0.0;
}

// cdist(train * w, test * w, 'sqeuclidean'): one thread per (test, train)
__global__ void k_py_wnd_dist(const double *train, const double *wts, const double *test, int n, int f, int ntest,
                              double *dist) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)n * ntest) return;
  const int q = e / n, i = e - (long)q * n;
  const double *tv = test + (long)q * f, *sv = train + (long)i * f;
  double ds = 0.0;
  for (int k = 0; k < f; k++) {
    const double wk = wts[k];
    const double diff = sv[k] * wk - tv[k] * wk;
    ds = ds + diff * diff;
  }
  dist[e] = ds;
}

// numpy's pairwise sum of at most 128 terms: fewer than 8 left to right,
// else eight running accumulators combined ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7))
// and the leftover terms
__device__ double np_pw_block(const double *a, long n) {
  if (n < 8) {
    double r = 0.0;
    for (long i = 0; i < n; i++) r += a[i];
    return r;
  }
  double r0 = a[0], r1 = a[1], r2 = a[2], r3 = a[3], r4 = a[4], r5 = a[5], r6 = a[6], r7 = a[7];
  long i = 8;
  for (; i < n - (n % 8); i += 8) {
    r0 += a[i]; r1 += a[i + 1]; r2 += a[i + 2]; r3 += a[i + 3];
    r4 += a[i + 4]; r5 += a[i + 5]; r6 += a[i + 6]; r7 += a[i + 7];
  }
  double res = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
  for (; i < n; i++) res += a[i];
  return res;
}

// numpy's pairwise sum: above 128 terms the halves (the first rounded down
// to a multiple of 8), summed depth first; the recursion runs on an
// explicit stack (n <= 8192 gives at most 7 levels)
__device__ double np_pairwise(const double *a, long n) {
  long off[16], len[16];
  double left[16];
  int stage[16];
  int sp = 0;
  off[0] = 0;
  len[0] = n;
  stage[0] = 0;
  double ret = 0.0;
  bool returning = false;
  while (true) {
    if (!returning) {
      if (len[sp] <= 128) {
        ret = np_pw_block(a + off[sp], len[sp]);
        returning = true;
      } else {
        long n2 = len[sp] / 2;
        n2 -= n2 % 8;
        stage[sp] = 0;
        off[sp + 1] = off[sp];
        len[sp + 1] = n2;
        sp++;
        continue;
      }
    }
    if (sp == 0) return ret;
    sp--;
    long n2 = len[sp] / 2;
    n2 -= n2 % 8;
    if (stage[sp] == 0) {  // left half done: evaluate the right half
      left[sp] = ret;
      stage[sp] = 1;
      off[sp + 1] = off[sp] + n2;
      len[sp + 1] = len[sp] - n2;
      sp++;
      returning = false;
    } else {  // both halves done
      ret = left[sp] + ret;
    }
  }
}

// np.add.reduce of a contiguous 1-D float64 array: pairwise within chunks of
// 8192 (the ufunc buffer), chunk sums added in order
__device__ double np_sum(const double *a, long n) {
  double res = 0.0;
  for (long s = 0; s < n; s += 8192) {
    const long m = n - s < 8192 ? n - s : 8192;
    const double p = np_pairwise(a + s, m);
    res = s == 0 ? p : res + p;
  }
  return res;
}

// per (test q, class c): the similarities d^-5 of the class's training rows
// (rows[start[c-1]..start[c]) in order) whose distance exceeds DBL_EPSILON,
// compacted into sims, and their numpy mean (NaN when there are none)
__global__ void k_py_wnd_means(const double *dist, const int *rows, const int *start, int n, int C, int ntest,
                               double *sims, double *cmean) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)ntest * C) return;
  const int q = e / C, c = e - (long)q * C;
  double *out = sims + (long)q * n + start[c];
  long k = 0;
  for (int j = start[c]; j < start[c + 1]; j++) {
    const double d = dist[(long)q * n + rows[j]];
    if (d > kDblEps) out[k++] = pow(d, -5.0);
  }
  cmean[e] = k ? np_sum(out, k) / (double)k : __longlong_as_double(0x7ff8000000000000LL);
}

// per test vector: no prediction if a class mean is NaN; else the
// normalisation factor (numpy sum of the class means), the marginal
// probabilities and the first argmax
__global__ void k_py_wnd_final(const double *cmean, int C, int ntest, double *prob, double *norm, long *pred) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ntest) return;
  const double *m = cmean + (long)q * C;
  bool any_nan = false;
  for (int c = 0; c < C; c++) any_nan |= m[c] != m[c];
  if (any_nan) {
    const double nan = __longlong_as_double(0x7ff8000000000000LL);
    for (int c = 0; c < C; c++) prob[(long)q * C + c] = nan;
    norm[q] = nan;
    pred[q] = 0;
    return;
  }
  const double nf = np_sum(m, C);
  long best = 0;
  double bp = 0;
  for (int c = 0; c < C; c++) {
    const double p = m[c] / nf;
    prob[(long)q * C + c] = p;
    if (c == 0 || p > bp) { bp = p; best = c + 1; }
  }
  norm[q] = nf;
  pred[q] = best;
}

void launch_py_normalize(double *M, int n, int f, int flags, double *mins, double *maxs, double *means,
                         double *stdevs, int *flag, hipStream_t s) {
  launch_fill_int(flag, 1, 0, s);
  const long total = (long)n * f;
  const int blocks = (int)std::min<long>(1024, (total + 255) / 256);
  k_py_any_invalid<<<blocks, 256, 0, s>>>(M, total, flag);
  k_py_norm<<<(f + 127) / 128, 128, 0, s>>>(M, n, f, flags, mins, maxs, means, stdevs, flag);
}
void launch_py_fisher(const double *data, const int *cls, int n, int f, int C, double *w, hipStream_t s) {
  k_py_fisher<<<(f + 127) / 128, 128, 0, s>>>(data, cls, n, f, C, w);
}
void launch_py_wnd5(const double *train, const double *w, const double *test, const int *rows, const int *start,
                    int n, int f, int C, int ntest, double *dist, double *sims, double *cmean, double *prob,
                    double *norm, long *pred, hipStream_t s) {
  const long e = (long)n * ntest;
  k_py_wnd_dist<<<(e + 255) / 256, 256, 0, s>>>(train, w, test, n, f, ntest, dist);
  const long ec = (long)ntest * C;
  k_py_wnd_means<<<(ec + 63) / 64, 64, 0, s>>>(dist, rows, start, n, C, ntest, sims, cmean);
  k_py_wnd_final<<<(ntest + 63) / 64, 64, 0, s>>>(cmean, C, ntest, prob, norm, pred);
}

}  // namespace wnd
