// Texture feature kernels: Haralick (co-occurrence statistics f1..f14),
// Tamura, Comb moments and Radon.  FP64.
//
// Haralick f14 needs the eigenvalue at index 2 of the real Schur form that
// the reference's Eigen 3.1.2 EigenSolver produces (CVIPtexture.cpp:993-1023),
// so Q, the Householder Hessenberg reduction and the Francis double-shift QR
// below perform the same floating-point operations in the same order as the
// CPU oracle's restatement of that algorithm: every reduction that feeds a
// decision is a per-thread sequential loop, parallelism is across the
// independent rows / columns a reflector touches.
#include "common.hpp"
#include "kernels.hpp"

namespace wnd {

// Memory fence for lanes of one wavefront exchanging values through LDS:
// a wave's LDS operations execute in issue order, so only compiler
// reordering has to be prevented.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- Eigen 3.1.2 evaluation orders (SSE2, two doubles per packet) used by
// the reference's f14 on row-major arrays (CVIPtexture.cpp:961-1023):
// eigen_redux is DenseBase::sum() over a contiguous range with linear
// vectorized traversal (Core/Redux.h:193-248); s is the index of the first
// 16-byte-aligned term.  x(k) supplies term k.
template <class F>
__device__ __forceinline__ double eigen_redux(int m, int s, F x) {
  if (s > m) s = m;
  const int as = ((m - s) / 2) * 2, as2 = ((m - s) / 4) * 4;
  const int end = s + as, end2 = s + as2;
  double res;
  if (as) {
    double a0 = x(s), a1 = x(s + 1);
    if (as > 2) {
      double b0 = x(s + 2), b1 = x(s + 3);
      for (int k = s + 4; k < end2; k += 4) {
        a0 = a0 + x(k); a1 = a1 + x(k + 1);
        b0 = b0 + x(k + 2); b1 = b1 + x(k + 3);
      }
      a0 = a0 + b0; a1 = a1 + b1;
      if (end > end2) { a0 = a0 + x(end2); a1 = a1 + x(end2 + 1); }
    }
    res = a0 + a1;
    for (int k = 0; k < s; k++) res = res + x(k);
    for (int k = end; k < m; k++) res = res + x(k);
  } else {
    res = x(0);
    for (int k = 1; k < m; k++) res = res + x(k);
  }
  return res;
}

// ===================================================================== Haralick
// quantisation (haralick.cpp:29-37) and tone presence
__global__ void k_har_quant(const Img *imgs, const Stats *st, const HarDesc *hd) {
  const Img im = imgs[blockIdx.y];
  const HarDesc H = hd[blockIdx.y];
  const Stats S = st[blockIdx.y];
  const long n = (long)im.h * im.w;
  const double scale255 = (255.0 / (S.mx - S.mn));
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double t = (im.p[i] - S.mn) * scale255;
    const int g = (t >= 0 && t < 256) ? (int)t : (t >= 256 ? 255 : 0);
    H.gray[i] = (uint8_t)g;
    H.lut[g] = 1;  // benign race: every writer stores 1
  }
}

// tone LUT (CVIPtexture.cpp:400-414): rank of each present gray level
__global__ void k_har_lut(const HarDesc *hd) {
  const HarDesc H = hd[blockIdx.x];
  if (threadIdx.x == 0) {
    int it = 0;
    for (int g = 0; g < 256; g++) H.lut[g] = H.lut[g] ? it++ : -1;
    *H.ng = it;
  }
}

// co-occurrence counts (CoOcMat_Angle_*, CVIPtexture.cpp:455-614) in LDS as the
// upper triangle of the symmetric matrix, then P = count / total
__device__ __forceinline__ int tri(int i, int j) { return i * (2 * 256 - i + 1) / 2 + (j - i); }

__global__ void __launch_bounds__(1024) k_har_cooc(const Img *imgs, const HarDesc *hd) {
  extern __shared__ uint32_t cnt[];  // 256*257/2
  __shared__ int lut[256];
  __shared__ int total;
  const int angle = blockIdx.x;
  const Img im = imgs[blockIdx.y];
  const HarDesc H = hd[blockIdx.y];
  const int ng = *H.ng;
  for (int e = threadIdx.x; e < 256 * 257 / 2; e += blockDim.x) cnt[e] = 0;
  if (threadIdx.x < 256) lut[threadIdx.x] = H.lut[threadIdx.x];
  if (threadIdx.x == 0) total = 0;
  __syncthreads();
  const int rows = im.h, cols = im.w;
  const int n = rows * cols;  // planes hold fewer than 2^31 pixels: 32-bit index arithmetic
  const gptr<const uint8_t> gray = to_global((const uint8_t *)H.gray);
  int local = 0;
  // four pixels per thread and pass: the pixel and neighbour loads are clamped
  // into the plane and issued together, the pairs outside it are not counted
  for (int i0 = threadIdx.x; i0 < n; i0 += 4 * blockDim.x) {
    int gv[4], g2v[4];
    bool okv[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int i = i0 + u * (int)blockDim.x;
      bool ok = i < n;
      const int ii = ok ? i : 0;
      const int row = ii / cols, col = ii - row * cols;
      int r2 = row, c2 = col;
      if (angle == 0) { c2 = col + 1; ok = ok && c2 < cols; }
      else if (angle == 1) { r2 = row + 1; c2 = col - 1; ok = ok && r2 < rows && c2 >= 0; }
      else if (angle == 2) { r2 = row + 1; ok = ok && r2 < rows; }
      else { r2 = row + 1; c2 = col + 1; ok = ok && r2 < rows && c2 < cols; }
      gv[u] = gray[ii];
      g2v[u] = gray[ok ? r2 * cols + c2 : ii];
      okv[u] = ok;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      if (!okv[u] || !gv[u] || !g2v[u]) continue;
      const int x = lut[gv[u]], y = lut[g2v[u]];
      if (x == y) atomicAdd(&cnt[tri(x, x)], 2u);
      else atomicAdd(&cnt[x < y ? tri(x, y) : tri(y, x)], 1u);
      local += 2;
    }
  }
  atomicAdd(&total, local);
  __syncthreads();
  const int count = total;
  double *P = H.P + (long)angle * ng * ng;
  for (int e = threadIdx.x; e < ng * ng; e += blockDim.x) {
    const int i = e / ng, j = e - i * ng;
    const uint32_t c = cnt[i <= j ? tri(i, j) : tri(j, i)];
    P[e] = count == 0 ? 0.0 : (double)c / count;
  }
  if (threadIdx.x == 0) H.count[angle] = count;
}

// f1..f13 (CVIPtexture.cpp:631-958) and Q (_Q_calc_eigen, 961-987) per angle
__global__ void __launch_bounds__(256) k_har_stats(const HarDesc *hd) {
  __shared__ double px[256], py[256], pxpy_s[512], pxpy_d[256];
  __shared__ double lse[512], lde[256], lhx[256], lhy[256];  // the entropy terms, formed in parallel
  __shared__ double shd[4];
  const int angle = blockIdx.x;
  const HarDesc H = hd[blockIdx.y];
  const int ng = *H.ng;
  const gptr<const double> P = to_global((const double *)H.P + (long)angle * ng * ng);
  const int t = threadIdx.x;
  const double LOG2 = log10(2.0), EPS = 0.000000001;
  // exact per-row / per-column marginals (sequential, as the reference
  // loops).  P is symmetric (k_har_cooc stores the counts as a triangle),
  // so row t equals column t term by term and both sums walk the column,
  // which the lanes read coalesced.
  // (each sequential sum loads its terms eight at a time, then adds them in
  // order: the loads are in flight together instead of one latency a term)
  if (t < ng) {
    double b = 0;
    int i = 0;
    for (; i + 8 <= ng; i += 8) {
      double v[8];
#pragma unroll
      for (int q = 0; q < 8; q++) v[q] = P[(i + q) * ng + t];
#pragma unroll
      for (int q = 0; q < 8; q++) b += v[q];
    }
    for (; i < ng; i++) b += P[i * ng + t];
    px[t] = b;
    py[t] = b;
  }
  // Pxpy[i+j] (anti-diagonals, raster order) and Pxpy[|i-j|] (f2, f10, f11)
  for (int k = t; k < 2 * ng - 1; k += blockDim.x) {
    double s = 0;
    const int i0 = k - (ng - 1) > 0 ? k - (ng - 1) : 0, i1 = k < ng - 1 ? k : ng - 1;
    int i = i0;
    for (; i + 8 <= i1 + 1; i += 8) {
      double v[8];
#pragma unroll
      for (int q = 0; q < 8; q++) v[q] = P[(i + q) * ng + (k - i - q)];
#pragma unroll
      for (int q = 0; q < 8; q++) s += v[q];
    }
    for (; i <= i1; i++) s += P[i * ng + (k - i)];
    pxpy_s[k] = s;
  }
  if (t < ng) {
    // both loads at clamped (in-range) columns every row, so they are not
    // behind a branch and can be in flight together; the adds are as before
    double s = 0;
    int i = 0;
    for (; i + 4 <= ng; i += 4) {
      double lo[4], hi[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        lo[q] = P[(i + q) * ng + max(i + q - t, 0)];
        hi[q] = P[(i + q) * ng + min(i + q + t, ng - 1)];
      }
#pragma unroll
      for (int q = 0; q < 4; q++) {
        if (i + q - t >= 0) s += lo[q];
        if (t > 0 && i + q + t < ng) s += hi[q];
      }
    }
    for (; i < ng; i++) {
      const double lo = P[i * ng + max(i - t, 0)], hi = P[i * ng + min(i + t, ng - 1)];
      if (i - t >= 0) s += lo;
      if (t > 0 && i + t < ng) s += hi;
    }
    pxpy_d[t] = s;
  }
  __syncthreads();
  // raster double sums (parallel)
  double f1 = 0, tmp3 = 0, mean4 = 0, idm = 0, ent = 0, hxy1 = 0, hxy2 = 0;
  const int nn = ng * ng, bs = blockDim.x;
  double pb[4] = {0.0, 0.0, 0.0, 0.0};  // this thread's next four elements, loaded ahead
  if (nn > 0)
#pragma unroll
    for (int q = 0; q < 4; q++) pb[q] = P[min(t + q * bs, nn - 1)];
  for (int e = t; e < nn; e += bs) {
    const int i = e / ng, j = e - i * ng;
    const double p = pb[0];
    pb[0] = pb[1];
    pb[1] = pb[2];
    pb[2] = pb[3];
    pb[3] = P[min(e + 4 * bs, nn - 1)];  // four elements ahead (clamped: unused past the end)
    f1 += p * p;
    tmp3 += i * j * p;
    mean4 += i * p;
    idm += p / (1 + (i - j) * (i - j));
    ent += p * log10(p + EPS) / LOG2;
    const double lg = log10(px[i] * py[j] + EPS);
    hxy1 -= p * lg / LOG2;
    hxy2 -= px[i] * py[j] * lg / LOG2;
  }
  f1 = block_reduce(f1, SumOp(), shd);
  tmp3 = block_reduce(tmp3, SumOp(), shd);
  mean4 = block_reduce(mean4, SumOp(), shd);
  idm = block_reduce(idm, SumOp(), shd);
  ent = block_reduce(ent, SumOp(), shd);
  hxy1 = block_reduce(hxy1, SumOp(), shd);
  hxy2 = block_reduce(hxy2, SumOp(), shd);
  double var4 = 0;
  {
    int e = t;
    for (; e + 7 * bs < nn; e += 8 * bs) {
      double v[8];
#pragma unroll
      for (int q = 0; q < 8; q++) v[q] = P[e + q * bs];
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const int i = (e + q * bs) / ng;
        var4 += (i - mean4) * (i - mean4) * v[q];
      }
    }
    for (; e < nn; e += bs) {
      const int i = e / ng;
      var4 += (i - mean4) * (i - mean4) * P[e];
    }
  }
  var4 = block_reduce(var4, SumOp(), shd);
  // the terms of the sum entropy, difference entropy and HX / HY, each formed
  // by one thread with the serial loop's operations; thread 0 then adds them
  // in the loops' order
  for (int k = t; k <= 2 * ng - 2; k += blockDim.x) {
    const double v = pxpy_s[k];
    lse[k] = v * log10(v + EPS) / LOG2;
  }
  if (t < ng) {
    lde[t] = pxpy_d[t] * log10(pxpy_d[t] + EPS) / LOG2;
    lhx[t] = px[t] * log10(px[t] + EPS) / LOG2;
    lhy[t] = py[t] * log10(py[t] + EPS) / LOG2;
  }
  __syncthreads();
  if (t == 0) {
    double *F = H.feat + angle * 14;
    F[0] = f1;
    double bigsum = 0;
    for (int n = 0; n < ng; n++) bigsum += n * n * pxpy_d[n];
    F[1] = bigsum;
    double meanx = 0, sum_sqrx = 0;
    for (int i = 0; i < ng; i++) { meanx += px[i] * i; sum_sqrx += px[i] * i * i; }
    const double sd = sqrt(sum_sqrx - (meanx * meanx));
    F[2] = (sd * sd == 0) ? 1.0 : (tmp3 - meanx * meanx) / (sd * sd);
    F[3] = var4;
    F[4] = idm;
    double savg = 0;
    for (int i = 0; i <= 2 * ng - 2; i++) savg += i * pxpy_s[i];
    F[5] = savg;
    double se = 0;
    for (int i = 2; i <= 2 * ng; i++) se -= lse[i - 2];  // i - 2 <= 2 ng - 2 throughout
    F[7] = se;
    double sv = 0;
    for (int i = 0; i <= 2 * ng - 2; i++) sv += (i - se) * (i - se) * pxpy_s[i];
    F[6] = sv;
    F[8] = -ent;
    double s = 0, ss = 0, de = 0;
    for (int i = 0; i < ng; i++) { s += i * pxpy_d[i]; ss += i * i * pxpy_d[i]; }
    F[9] = ss - s * s;
    for (int i = 0; i < ng; i++) de += lde[i];
    F[10] = -de;
    double hx = 0, hy = 0;
    for (int i = 0; i < ng; i++) {
      hx -= lhx[i];
      hy -= lhy[i];
    }
    const double hxy = -ent;
    const double mxh = hx > hy ? hx : hy;
    F[11] = (mxh == 0) ? 1.0 : (hxy - hxy1) / mxh;
    F[12] = sqrt(fabs(1 - exp(-2.0 * (hxy2 - hxy))));
  }
  // marginals of _Q_calc_eigen: px = P.rowwise().sum() (row t of the
  // row-major array starts at linear index t*ng; its terms are read down
  // column t, equal by symmetry), py = P.colwise().sum() (strided: a
  // running sum, equal to py above)
  if (t < ng) {
    H.marg[angle * 512 + t] = eigen_redux(ng, (int)(((long)t * ng) & 1), [&](int k) { return P[(long)k * ng + t]; });
    H.marg[angle * 512 + 256 + t] = py[t];
  }
}

// Q (_Q_calc_eigen, CVIPtexture.cpp:961-987):
//   Q(i,j) = (P.row(i)*P.row(j) / (px(i) py + 1e-13)).sum(),
// an Eigen redux over an aligned expression: terms k < 4*floor(Ng/4) go to
// four accumulators by k mod 4, which are paired (0+2, 1+3); the next two
// terms (if Ng mod 4 >= 2) join the pairs, the pairs are added, and an odd
// last term comes last.  Every term is >= +0, so an accumulator that
// starts at +0 and skips zero terms holds the same bits: each row visits
// only the k with P(i,k) != 0 (compacted per wavefront), and skips
// P(j,k) == 0.  P is symmetric (pairs are counted both ways), so P(j,k) is
// read as P(k,j): along row k, coalesced over j.
constexpr int QRB = 4;  // row blocks per angle
struct Redux4 {
  double a0, a1, a2, a3, e0, e1, tl;
};
// k is the same for every lane of a wavefront, so the branches are uniform
__device__ __forceinline__ void r4_add(Redux4 &r, int k, int as2, int as, double term) {
  if (k < as2 || (as2 == 0 && k < as)) {  // four accumulators; ng 2..3: the single packet (x0, x1)
    switch (k & 3) {
      case 0: r.a0 += term; break;
      case 1: r.a1 += term; break;
      case 2: r.a2 += term; break;
      default: r.a3 += term; break;
    }
  } else if (k < as) {
    if (k == as2) r.e0 += term;
    else r.e1 += term;
  } else {
    r.tl += term;  // odd last term (or x0 when ng == 1)
  }
}
__device__ __forceinline__ double r4_finish(const Redux4 &r, int ng) {
  if (ng == 1) return r.tl;
  double res;
  if (ng >= 4) {
    double l0 = r.a0 + r.a2, l1 = r.a1 + r.a3;
    if ((ng / 2) * 2 > (ng / 4) * 4) { l0 = l0 + r.e0; l1 = l1 + r.e1; }
    res = l0 + l1;
  } else {
    res = r.a0 + r.a1;
  }
  if (ng & 1) res = res + r.tl;
  return res;
}

__global__ void __launch_bounds__(256) k_har_q(const HarDesc *hd, int nimg) {
  __shared__ int lk[4][256];
  __shared__ double la[4][256], ld[4][256];
  int item, rb;
  if (!xcd_item(QRB, 4 * nimg, item, rb)) return;  // the row blocks of one P share an XCD
  const int angle = item & 3;
  const HarDesc H = hd[item >> 2];
  const int ng = *H.ng;
  const double *P = H.P + (long)angle * ng * ng;
  const double *px = H.marg + angle * 512, *py = px + 256;
  double *Q = H.Q + (long)angle * ng * ng;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int i = rb * 4 + w; i < ng; i += QRB * 4) {
    const double pxi = px[i];
    int cnt = 0;
    for (int k0 = 0; k0 < ng; k0 += 64) {
      const int k = k0 + lane;
      const double a = k < ng ? to_global(P)[(long)i * ng + k] : 0.0;
      const unsigned long long m = __ballot(a != 0.0);
      if (a != 0.0) {
        const int pos = cnt + __popcll(m & ((1ull << lane) - 1));
        lk[w][pos] = k;
        la[w][pos] = a;
        ld[w][pos] = py[k] * pxi + 0.0000000000001;
      }
      cnt += __popcll(m);
    }
    wave_sync();
    // four output columns per lane, one pass over the row's list
    const int as2 = (ng / 4) * 4, as = (ng / 2) * 2;
    Redux4 r[4];
#pragma unroll
    for (int q = 0; q < 4; q++) r[q] = Redux4{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    // list entries in order, two at a time so that 8 row loads are in flight
    const gptr<const double> Pg = to_global(P);
    int kk = 0;
    for (; kk + 2 <= cnt; kk += 2) {
      int k[2];
      double a[2], d[2], b[2][4];
#pragma unroll
      for (int u = 0; u < 2; u++) {
        k[u] = lk[w][kk + u];
        a[u] = la[w][kk + u];
        d[u] = ld[w][kk + u];
      }
#pragma unroll
      for (int u = 0; u < 2; u++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int col = lane + 64 * q;
          const double v = Pg[(unsigned)(k[u] * ng + min(col, ng - 1))];
          b[u][q] = col < ng ? v : 0.0;
        }
#pragma unroll
      for (int u = 0; u < 2; u++)
#pragma unroll
        for (int q = 0; q < 4; q++)
          if (b[u][q] != 0.0) r4_add(r[q], k[u], as2, as, (a[u] * b[u][q]) / d[u]);
    }
    for (; kk < cnt; kk++) {
      const int k = lk[w][kk];
      const double a = la[w][kk], d = ld[w][kk];
      double b[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int col = lane + 64 * q;
        const double v = Pg[(unsigned)(k * ng + min(col, ng - 1))];
        b[q] = col < ng ? v : 0.0;
      }
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (b[q] != 0.0) r4_add(r[q], k, as2, as, (a * b[q]) / d);
    }
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (lane + 64 * q < ng) Q[(long)i * ng + lane + 64 * q] = r4_finish(r[q], ng);
    wave_sync();  // the lists are rebuilt for the next row
  }
}

// --------------------------------------------------------- eigenvalues of Q
// Householder vector (Eigen makeHouseholder) of x[0..m-1], one thread.
__device__ void make_hh(const double *x, int m, double *ess, double *tau, double *beta) {
  double tail = 0;
  for (int i = 1; i < m; i++) tail += x[i] * x[i];
  const double c0 = x[0];
  if (tail == 0.0) {
    *tau = 0; *beta = c0;
    for (int i = 1; i < m; i++) ess[i - 1] = 0;
    return;
  }
  double b = sqrt(c0 * c0 + tail);
  if (c0 >= 0) b = -b;
  for (int i = 1; i < m; i++) ess[i - 1] = x[i] / (c0 - b);
  *tau = (b - c0) / b;
  *beta = b;
}

// One block per (angle, image): Hessenberg reduction of Q (n x n, row-major,
// in place), then (k_har_qr) the Francis QR on the active windows.  Only the
// values of index 0..2 of the Schur diagonal are consumed (f14 = sqrt of
// index 2 when that eigenvalue is real).
constexpr int QB = 28;       // bulge-chase steps per LDS window
constexpr int QW = QB + 4;   // window side (<= 64: one wavefront lane per row/column)
constexpr int QWP = QW + 1;  // padded LDS row of the window
// window and strip tiles are loaded as two rows of 32 lanes per wave instruction
static_assert(QW <= 32 && QW % 2 == 0 && QB + 2 <= QW, "QR window tiles must fit 32-lane rows");
static_assert(QW * QWP >= 256, "the window buffer also holds the n <= 256 row sums of the Schur norm");
constexpr int HT = 32;       // Hessenberg: columns per row-dot tile (n x HT in LDS)
constexpr int HTP = HT + 1;  // padded tile row (bank-conflict free row reads)
constexpr int HS = 256 / HT; // row stride of the tile-loading threads; each loads HT rows
constexpr int HR = 16;       // rows in flight per thread in the column pass (a multiple of 4)

// Reflector of column s below the diagonal (Eigen makeHouseholderInPlace),
// x[0..rm-1] in LDS: one thread's sequential squared norm of the tail (the
// column is strided in row-major storage, so Eigen sums it in order), then
// the essential part element-parallel.  Writes beta and the zeros of the
// Hessenberg form into column s, and the reflector into row s of `store`
// (essential part, tau in column n-1) for the rows above the active block.
// Returns tau.
__device__ __forceinline__ double hess_reflector(gptr<double> A, int n, int s, int rm, const double *x, double *ess,
                                                 double *sc, gptr<double> store) {
  const int t = threadIdx.x;
  if (t == 0) {
    double tail = 0;
    for (int k = 1; k < rm; k++) tail += x[k] * x[k];
    const double c0 = x[0];
    if (tail == 0.0) {
      sc[0] = 0; sc[1] = c0; sc[2] = 0; sc[3] = 0;
    } else {
      double b = sqrt(c0 * c0 + tail);
      if (c0 >= 0) b = -b;
      sc[0] = (b - c0) / b;
      sc[1] = b;
      sc[2] = c0 - b;
      sc[3] = 1;
    }
  }
  __syncthreads();
  const double tau = sc[0];
  const bool nz = sc[3] != 0;
  const double den = sc[2];
  for (int k = 1 + t; k < rm; k += blockDim.x) {
    const double e = nz ? x[k] / den : 0.0;
    ess[k - 1] = e;
    store[(long)s * n + k - 1] = e;
  }
  if (t == 0) {
    A[(long)(s + 1) * n + s] = sc[1];
    store[(long)s * n + n - 1] = tau;
  }
  for (int k = 1 + t; k < rm; k += blockDim.x) A[(long)(s + 1 + k) * n + s] = 0.0;  // matrixH(): Hessenberg part only
  __syncthreads();
  return tau;
}

// Column pass after step i (columns c = i+2..n-1, one thread per column;
// threads left over split the rows above the dot region):
//   upd  A = H A and A = A H of step i on rows i+1..n-1 (the rows above are
//        deferred).  The tile pass computed H A only in LDS, so the left
//        update is applied here to the stored values first, with the same
//        operations: A(r,c) -= fac(r) * twold[c] (fac = tau for row i+1,
//        tau * ess[r-i-2] below; twold = the left dot of step i), then
//        A(r,c) -= tu[r] * ess[c-i-2], with
//        tu[r] = tau * (row dot), Eigen's (tau * tmp) * essential^T
//   dot  the left dot of step i+1 on the updated column, in Eigen's ColMajor
//        GEMV order on bottom^T (general_matrix_vector_product<ColMajor>,
//        GeneralMatrixVector.h:31-285): inner terms (rows i+3..n-1 times nx)
//        in groups of four (g, g+o1, g+2, g+o3), added pairwise for outputs
//        in the packet range and left to right for an odd last output;
//        leftover terms one by one; a skipped leading term last; then
//        + A(i+2, c) (applyHouseholderOnTheLeft's tmp += row(0)).  -> tw[c]
__device__ __forceinline__ void hess_colpass(gptr<double> A, int n, int i, bool upd, const double *ess,
                                             const double *tu, double tau, const double *twold, bool dot,
                                             const double *nx, double *tw) {
  const int t = threadIdx.x;
  const int mB = n - i - 2;
  if (mB <= 0) return;
  const int k = min(8, (int)blockDim.x / mB);
  const int cc = t % mB, sub = t / mB;
  if (sub >= k) return;
  const int c = i + 2 + cc;
  const double f = upd ? ess[cc] : 0.0;
  const double wl = upd ? twold[c] : 0.0;  // left dot of step i for this column
  // left factor of row r (r >= i+1) of step i: tau * (1, ess)
  auto lfac = [&](int r) { return r == i + 1 ? tau : tau * ess[r - i - 2]; };
  const int d0 = dot ? i + 2 : n;  // first row of the dot region
  if (upd && !(dot && k > 1 && sub == 0)) {
    const int parts = (dot && k > 1) ? k - 1 : k, ps = (dot && k > 1) ? sub - 1 : sub;
    const int base = i + 1;  // upd implies i >= 0
    const int ra = base + (d0 - base) * ps / parts, rb = base + (d0 - base) * (ps + 1) / parts;
    unsigned o = (unsigned)(ra * n + c);
    int r = ra;
    for (; r + HR <= rb; r += HR, o += HR * n) {
      double v[HR];
#pragma unroll
      for (int q = 0; q < HR; q++) v[q] = A[o + q * n];
#pragma unroll
      for (int q = 0; q < HR; q++) A[o + q * n] = (v[q] - lfac(r + q) * wl) - tu[r + q] * f;
    }
    for (; r < rb; r++, o += n) A[o] = (A[o] - lfac(r) * wl) - tu[r] * f;
  }
  if (!dot || sub != 0) return;
  const int s = i + 1;       // the step whose left dot this is
  const int ms = n - s - 2;  // its inner terms: rows s+2..n-1
  const bool odd_n = n & 1;
  const long L = (long)(s + 2) * n + (s + 1);  // bottom^T(0,0) in row-major storage
  const int skip = ((L & 1) && odd_n && ms > 0) ? 1 : 0;
  const bool packet = cc < ((n - s - 1) & ~1);
  const int bound = ((ms - skip) / 4) * 4 + skip;
  const unsigned ot = (unsigned)((s + 1) * n + c);
  double top = A[ot];
  if (upd) { top = (top - lfac(s + 1) * wl) - tu[s + 1] * f; A[ot] = top; }
  // rows s+2+j are addressed as A[o0 + j*n]: a uniform base and a 32-bit
  // offset per load (one register per address)
  const unsigned o0 = ot + n;  // row s+2: term 0
  double x0 = 0.0;
  if (skip) {
    double v = A[o0];
    if (upd) { v = (v - lfac(s + 2) * wl) - tu[s + 2] * f; A[o0] = v; }
    x0 = v * nx[0];
  }
  double acc = 0.0;
  int g = skip;
  for (; g + HR <= bound; g += HR) {
    double v[HR];
#pragma unroll
    for (int q = 0; q < HR; q++) v[q] = A[o0 + (g + q) * n];
    if (upd) {
#pragma unroll
      for (int q = 0; q < HR; q++) {
        v[q] = (v[q] - lfac(s + 2 + g + q) * wl) - tu[s + 2 + g + q] * f;
        A[o0 + (g + q) * n] = v[q];
      }
    }
#pragma unroll
    for (int h = 0; h < HR; h += 4) {
      const double a = v[h] * nx[g + h], p1 = v[h + 1] * nx[g + h + 1];
      const double c2 = v[h + 2] * nx[g + h + 2], p3 = v[h + 3] * nx[g + h + 3];
      const double b = odd_n ? p3 : p1, d = odd_n ? p1 : p3;
      if (packet) acc = acc + ((a + b) + (c2 + d));
      else acc = (((acc + a) + b) + c2) + d;
    }
  }
  for (; g < bound; g += 4) {
    double v[4];
#pragma unroll
    for (int q = 0; q < 4; q++) v[q] = A[o0 + (g + q) * n];
    if (upd) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        v[q] = (v[q] - lfac(s + 2 + g + q) * wl) - tu[s + 2 + g + q] * f;
        A[o0 + (g + q) * n] = v[q];
      }
    }
    const double a = v[0] * nx[g], p1 = v[1] * nx[g + 1], c2 = v[2] * nx[g + 2], p3 = v[3] * nx[g + 3];
    const double b = odd_n ? p3 : p1, d = odd_n ? p1 : p3;
    if (packet) acc = acc + ((a + b) + (c2 + d));
    else acc = (((acc + a) + b) + c2) + d;
  }
  for (int j = bound; j < ms; j++) {
    double v = A[o0 + j * n];
    if (upd) { v = (v - lfac(s + 2 + j) * wl) - tu[s + 2 + j] * f; A[o0 + j * n] = v; }
    acc = acc + v * nx[j];
  }
  if (skip) acc = acc + x0;
  tw[c] = acc + top;
}

// One block per (angle, image): HessenbergDecomposition::_compute
// (Eigenvalues/HessenbergDecomposition.h:292-318) of Q (n x n, row-major, in
// place).  Rows 0..i only ever receive the right reflectors of steps >= i,
// and nothing else reads them, so step i works on the active rows
// i+1..n-1 and the rows above are brought up to date at the end (deferred
// pass); every element still receives the same operations in the same
// order.  Two passes over the active block per step:
//   tile pass    columns i+1..n-1 in HT-column LDS tiles of the active rows.
//                The loading thread applies A = H A of step i (row i+1 -=
//                tau*tw, bottom -= (tau*ess)*tw; tw = the left dot from the
//                previous column pass) in LDS only (the column pass applies
//                it again to the stored values, so the block is written once
//                per step); the row's thread
//                then accumulates the row dot in Eigen's RowMajor GEMV order
//                (general_matrix_vector_product<RowMajor>,
//                GeneralMatrixVector.h:289-552): even and odd terms in two
//                lanes, the lanes added, an odd last term, + A(r, i+1).
//   column i+1   A = A H on column i+1 (factor 1), then the reflector of
//                step i+1 from it (kept in the co-occurrence buffer, free
//                once Q is built, for the deferred pass).
//   column pass  A = A H on columns i+2..n-1 fused with the left dot of
//                step i+1 (hess_colpass).
// Eigen's last step (one row below the diagonal) scales by 1 - tau with
// tau = 0 (a one-element reflector): an exact no-op, not performed.
// Two blocks per CU (waves_per_eu 2): the ~500 matrices in flight then fit
// in the MALL, which serves the re-reads of later steps.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_har_hess(const HarDesc *hd, const int *order) {
  __shared__ double tile[256 * HTP];
  __shared__ double essb[2][256];  // reflector of step i and of step i+1
  __shared__ double twb[2][256];   // left dots by column: current step, next step
  __shared__ double tu[256];       // tau * right dot of the current step, by row
  __shared__ double xcol[256];
  __shared__ double sc[4];
  const int angle = blockIdx.x;
  const HarDesc H = hd[order[blockIdx.y]];
  const int n = *H.ng;
  if (n < 2) return;
  gptr<double> A = to_global(H.Q + (long)angle * n * n);
  gptr<double> Rf = to_global(H.P + (long)angle * n * n);  // reflector store
  const int t = threadIdx.x;
  for (int k = t; k < n - 1; k += blockDim.x) xcol[k] = A[(long)(1 + k) * n];
  __syncthreads();
  double tau = hess_reflector(A, n, 0, n - 1, xcol, essb[0], sc, Rf);
  if (n - 1 >= 2) hess_colpass(A, n, -1, false, nullptr, nullptr, 0.0, nullptr, true, essb[0], twb[0]);
  __syncthreads();
  int pp = 0;
  const int lc = t % HT, lr = t / HT;
  for (int i = 0; i + 2 < n; ++i) {
    const int rem = n - i - 1, m = rem - 1, asz = m & ~1;
    const int rb = i + 1;  // first active row; tile row k holds row rb + k
    const double *ess = essb[pp];
    const double *tw = twb[pp];
    double rv[HT];
    auto load = [&](int ct) {
      if (ct + lc < rem) {
        int r0 = lr;
        asm volatile("" : "+v"(r0));  // per-row values are formed at use, not hoisted out of the column loop
        const int c = i + 1 + ct + lc;
        // rows past the end re-read row n-1: branch-free, so all HT loads issue back to back
#pragma unroll
        for (int q = 0; q < HT; q++) rv[q] = A[(unsigned)(min(rb + r0 + HS * q, n - 1) * n + c)];
      }
    };
    double l0 = 0.0, l1 = 0.0, tl = 0.0, col0 = 0.0;
    load(0);
    for (int ct = 0; ct < rem; ct += HT) {
      const int ncol = (rem - ct < HT) ? rem - ct : HT;
      if (lc < ncol) {
        const int c = i + 1 + ct + lc;
        const double w = tw[c];
        int r0 = lr;
        asm volatile("" : "+v"(r0));
        double *tl0 = tile + (r0 * HTP + lc);
        // every LDS operand is read before the arithmetic (clamped indices,
        // unused where clamped), so the reads are in flight together rather
        // than one round trip per row
        double e[HT];
#pragma unroll
        for (int q = 0; q < HT; q++) e[q] = ess[min(max(rb + r0 + HS * q - i - 2, 0), 255)];
#pragma unroll
        for (int q = 0; q < HT; q++) {
          const int r = rb + r0 + HS * q;
          const double fac = (r == i + 1) ? tau : tau * e[q];
          tl0[HS * q * HTP] = (r < n) ? rv[q] - fac * w : rv[q];  // rows >= n are never read
        }
      }
      __syncthreads();
      if (ct + HT < rem) load(ct + HT);  // next tile's loads in flight during the row sums
      if (t < rem) {
        const double *tr = tile + t * HTP;
#pragma unroll
        for (int h = 0; h < HT; h += HT / 2) {  // half a tile row at a time: its LDS reads in flight together
          double tv[HT / 2], te[HT / 2];
#pragma unroll
          for (int u = 0; u < HT / 2; u++) {
            tv[u] = tr[h + u];
            te[u] = ess[min(max(ct + h + u - 1, 0), 255)];
          }
#pragma unroll
          for (int u = 0; u < HT / 2; u++) {
            const int cc = h + u;
            if (cc < ncol) {
              const int j = ct + cc - 1;  // ct is even: j is even for odd cc
              const double v = tv[u];
              if (j < 0) col0 = v;
              else if (j < asz) {
                if (cc & 1) l0 = l0 + v * te[u];
                else l1 = l1 + v * te[u];
              } else tl = v * te[u];
            }
          }
        }
      }
      __syncthreads();
    }
    if (t < rem) {
      const int r = rb + t;
      double tmp = 0.0;
      if (asz > 0) tmp = tmp + (l0 + l1);
      if (m & 1) tmp = tmp + tl;
      const double tt = (0.0 + tmp) + col0;
      const double tuv = tau * tt;
      tu[r] = tuv;
      const double v = col0 - tuv;  // column i+1 of A = A H: (tau*tmp) * 1
      A[(long)r * n + i + 1] = v;
      if (r >= i + 2) xcol[r - i - 2] = v;
    }
    __syncthreads();
    const double tau_n = hess_reflector(A, n, i + 1, m, xcol, essb[pp ^ 1], sc, Rf);
    hess_colpass(A, n, i, true, ess, tu, tau, tw, m >= 2, essb[pp ^ 1], twb[pp ^ 1]);
    __syncthreads();
    pp ^= 1;
    tau = tau_n;
  }
  // Deferred pass: row r (r <= n-3) receives the right reflectors of steps
  // r..n-3 in order, exactly as the column passes would have applied them:
  // row dot in the RowMajor GEMV order (lanes l0 / l1 of one row are the
  // two threads of a pair), tu = tau * (dot + A(r, i+1)), A(r, i+1) -= tu,
  // A(r, c) -= tu * ess[c-i-2].  Chunks of rows r0.. are staged in LDS with
  // their live columns r0+1..n-1 only, as many rows as the tile buffer
  // holds (at most 128); reflectors stream through a two-deep register
  // pipeline (one element per thread) into two LDS buffers.
  double *rows = tile;
  const int lane = t & 63, wv = t >> 6;
  for (int r0 = 0; r0 < n - 2;) {
    const int cb = r0 + 1, wd = n - cb, wp = wd | 1;  // staged columns cb..n-1, odd row stride
    int nr = (256 * HTP) / wp;
    if (nr > 128) nr = 128;
    if (nr > n - 2 - r0) nr = n - 2 - r0;
    for (int rr = wv; rr < nr; rr += 4)
      for (int c = lane; c < wd; c += 64) rows[rr * wp + c] = A[(long)(r0 + rr) * n + cb + c];
    // reflector r0 in LDS buffer 0; reflectors r0+1..r0+4 in flight in a
    // ring of four register slots, slot k holding reflector r0+s with
    // s = k (mod 4): the step loop is unrolled by four so that every slot
    // is named statically and consumed four steps after its load
    auto fetch = [&](int i, double &e, double &tv) {
      e = 0.0;
      tv = 0.0;
      if (i <= n - 3) {
        if (t < n - i - 2) e = Rf[(long)i * n + t];
        tv = Rf[(long)i * n + n - 1];
      }
    };
    double e0, t0, e1, t1, e2, t2, e3, t3;
    fetch(r0, e0, t0);
    essb[0][t] = e0;
    double ti = t0;
    fetch(r0 + 1, e1, t1);
    fetch(r0 + 2, e2, t2);
    fetch(r0 + 3, e3, t3);
    fetch(r0 + 4, e0, t0);
    __syncthreads();
    // step i = r0 + s; afterwards reflector i+1 (slot (s+1) mod 4) goes to
    // LDS and that slot is refilled with reflector i+5
    auto step = [&](int s, double &en, double &tn) {
      const int i = r0 + s;
      const double *eb = essb[s & 1];
      const int m = n - i - 2, asz = m & ~1;
      const int na = (i + 1 - r0 < nr) ? i + 1 - r0 : nr;  // rows r0..i of the chunk take part
      if (t < 2 * na) {
        const int rr = t >> 1, par = t & 1;
        const double *row = rows + rr * wp + (i + 2 - cb);
        double l = 0.0;
#pragma unroll 8
        for (int j = par; j < asz; j += 2) l = l + row[j] * eb[j];
        const double l1 = __shfl_xor(l, 1);
        if (par == 0) {
          double tmp = 0.0;
          if (asz > 0) tmp = tmp + (l + l1);
          if (m & 1) tmp = tmp + row[m - 1] * eb[m - 1];
          tu[rr] = ti * ((0.0 + tmp) + rows[rr * wp + (i + 1 - cb)]);
        }
      }
      __syncthreads();
      // column c (c = 0: column i+1, factor 1) of rows wv, wv+4, ..: the
      // rows' entries are independent, so their LDS accesses overlap
      for (int c = lane; c <= m; c += 64) {
        double *col = rows + (i + 1 - cb) + c;
        if (c == 0) {
#pragma unroll 4
          for (int rr = wv; rr < na; rr += 4) col[rr * wp] = col[rr * wp] - tu[rr];
        } else {
          const double e = eb[c - 1];
#pragma unroll 4
          for (int rr = wv; rr < na; rr += 4) col[rr * wp] = col[rr * wp] - tu[rr] * e;
        }
      }
      essb[(s + 1) & 1][t] = en;
      ti = tn;
      fetch(i + 5, en, tn);
      __syncthreads();
    };
    for (int s = 0; r0 + s <= n - 3; s += 4) {
      step(s, e1, t1);
      if (r0 + s + 1 > n - 3) break;
      step(s + 1, e2, t2);
      if (r0 + s + 2 > n - 3) break;
      step(s + 2, e3, t3);
      if (r0 + s + 3 > n - 3) break;
      step(s + 3, e0, t0);
    }
    for (int rr = wv; rr < nr; rr += 4)
      for (int c = lane; c < wd; c += 64) A[(long)(r0 + rr) * n + cb + c] = rows[rr * wp + c];
    __syncthreads();
    r0 += nr;
  }
}

// Francis QR on the Hessenberg matrix left in H.Q by k_har_hess.
// NT threads per (angle, image).  The Schur iteration is a chain of
// dependent steps, so for large batches many matrices per CU (NT = 64: a
// small LDS window and one wave each) beat many waves per matrix.  For a
// few matrices (a call of one or a few images) NT = 256: the chase still
// runs on wave 0, while the strip updates after each window, whose columns
// and rows are independent, take one round over 256 threads instead of
// several over 64 (and the rows above the window one LDS chunk of up to 256
// rows instead of chunks of 32).  Every entry receives the same operations
// in the same order either way.
template <int NT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT == 64 ? 4 : 1)))
k_har_qr(const HarDesc *hd, const int *order) {
  constexpr int RS = NT / 32;             // window rows per pass of the block (32 columns per row)
  constexpr int CH = NT == 64 ? QW : NT;  // rows above the window staged per chunk
  __shared__ double Rw[CH * QWP];
  __shared__ double rf[QB][4];
  __shared__ double sc[8];
  __shared__ int si[8];
  __shared__ int ired[8];
  __shared__ double dred[8];
  __shared__ int s_iu;
  const int angle = blockIdx.x;
  const HarDesc H = hd[order[blockIdx.y]];
  const int n = *H.ng;
  gptr<double> A = to_global(H.Q + (long)angle * n * n);
  const int t = threadIdx.x;
#define AT(r, c) A[(long)(r) * n + (c)]
  // ---- real Schur form (RealSchur::compute without U); left updates are
  // restricted to columns <= iu (entries right of the active window are
  // never read again), right updates cover rows 0..min(iu,k+3) as in Eigen.
  // RealSchur::computeNormOfT (RealSchur.h:274-285) on row-major storage:
  // each row segment T(j, max(j-1,0)..n-1) is a contiguous Eigen redux, and
  // the segment sums are added row by row
  double *segs = Rw;  // the window buffer is free until the first window
  for (int j = t; j < n; j += blockDim.x) {
    const int k0 = j - 1 > 0 ? j - 1 : 0;
    gptr<const double> row = A + ((long)j * n + k0);
    segs[j] = eigen_redux(n - k0, (int)(((long)j * n + k0) & 1), [&](int k) { return fabs(row[k]); });
  }
  __syncthreads();
  if (t == 0) {
    double sn = 0;
    for (int j = 0; j < n; j++) sn += segs[j];
    dred[0] = sn;
  }
  __syncthreads();
  const double norm = dred[0];
  const double eps = kDblEps;
  // decision state (iu, iteration counts, exceptional shift) lives in
  // thread 0; the block follows through s_iu / si / sc
  int iu = n - 1, iter = 0, totalIter = 0;
  double exshift = 0;
  double work = 0;  // thread 0: flops of the updates issued (12 per 3-vector, 6 per 2-vector / rotation)
  if (t == 0) s_iu = n - 1;
  __syncthreads();
  while (norm != 0) {
    __syncthreads();  // every thread has consumed the previous decision
    const int iub = s_iu;
    if (iub < 0) break;
    // findSmallSubdiagEntry: the largest res in (0, iu] whose subdiagonal
    // entry is negligible, else 0 -- every candidate tested at once
    // (a lane's candidates are res = iub - t - 64k, k < 4 since n <= 256: all
    // their entries are loaded at once, then tested in descending order)
    int cand = 0;
    {
      double dd0[4] = {0, 0, 0, 0}, dd1[4] = {0, 0, 0, 0}, sd[4] = {0, 0, 0, 0};
      if (n >= 2)  // otherwise there is no candidate
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const int res = min(max(iub - t - 64 * k, 1), n - 1);  // in range; unused when res <= 0
          dd0[k] = AT(res - 1, res - 1);
          dd1[k] = AT(res, res);
          sd[k] = AT(res, res - 1);
        }
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int res = iub - t - 64 * k;
        if (res > 0 && cand == 0) {
          double s = fabs(dd0[k]) + fabs(dd1[k]);
          if (s == 0.0) s = norm;
          if (fabs(sd[k]) < eps * s) cand = res;
        }
      }
    }
    const int il = block_reduce(cand, MaxOp(), ired);
    if (t == 0) {
      int action = 3;
      if (il == iu) {  // one root
        AT(iu, iu) = AT(iu, iu) + exshift;
        if (iu > 0) AT(iu, iu - 1) = 0;
        iu--; iter = 0;
        action = 0;
      } else if (il == iu - 1) {  // two roots (RealSchur::splitOffTwoRows)
        action = 1;
        const double p = 0.5 * (AT(iu - 1, iu - 1) - AT(iu, iu));
        const double q = p * p + AT(iu, iu - 1) * AT(iu - 1, iu);
        AT(iu, iu) += exshift;
        AT(iu - 1, iu - 1) += exshift;
        si[3] = q >= 0;
        if (q >= 0) {
          const double z = sqrt(fabs(q));
          const double pp = (p >= 0) ? p + z : p - z, qq = AT(iu, iu - 1);
          double c, s;  // JacobiRotation::makeGivens, real case
          if (qq == 0) { c = pp < 0 ? -1 : 1; s = 0; }
          else if (pp == 0) { c = 0; s = qq < 0 ? 1 : -1; }
          else if (fabs(pp) > fabs(qq)) {
            const double tt = qq / pp;
            double u = sqrt(1 + tt * tt);
            if (pp < 0) u = -u;
            c = 1 / u; s = -tt * c;
          } else {
            const double tt = pp / qq;
            double u = sqrt(1 + tt * tt);
            if (qq < 0) u = -u;
            s = -1 / u; c = -tt * s;
          }
          sc[1] = c; sc[2] = s;
          for (int j = iu - 1; j <= iu; j++) {  // rot.adjoint() on the left, columns <= iu
            const double xv = AT(iu - 1, j), yv = AT(iu, j);
            AT(iu - 1, j) = c * xv - s * yv;
            AT(iu, j) = s * xv + c * yv;
          }
        }
        si[1] = iu;  // rows 0..iu get the right rotation
        if (q >= 0) work += 6.0 * (2 + iu + 1);
        iu -= 2; iter = 0;
      } else {  // Francis QR step (RealSchur::computeShift)
        double sh0 = AT(iu, iu), sh1 = AT(iu - 1, iu - 1), sh2 = AT(iu, iu - 1) * AT(iu - 1, iu);
        double shift_amt = 0;
        int shift_kind = 0;
        if (iter == 10) {
          exshift += sh0;
          shift_kind = 1; shift_amt = sh0;
          const double s = fabs(AT(iu, iu - 1)) + fabs(AT(iu - 1, iu - 2));
          sh0 = 0.75 * s; sh1 = 0.75 * s; sh2 = -0.4375 * s * s;
        }
        if (iter == 30) {
          double s = (sh1 - sh0) / 2.0;
          s = s * s + sh2;
          if (s > 0) {
            s = sqrt(s);
            if (sh1 < sh0) s = -s;
            s = s + (sh1 - sh0) / 2.0;
            s = sh0 - sh2 / s;
            exshift += s;
            shift_kind = 2; shift_amt = s;
            sh0 = sh1 = sh2 = 0.964;
          }
        }
        iter++; totalIter++;
        if (totalIter > 40 * n) action = 3;
        else {
          action = 2;
          if (shift_kind) work += iu + 1;
          si[4] = shift_kind;
          sc[3] = shift_amt;
          sc[4] = sh0; sc[5] = sh1; sc[6] = sh2;
          si[5] = iu;
        }
      }
      si[0] = action;
      s_iu = action == 3 ? -1 : iu;
    }
    __syncthreads();
    const int action = si[0];
    if (action == 3) break;
    if (action == 0) continue;
    if (action == 1) {
      const int iu2 = si[1];
      if (si[3]) {
        const double c = sc[1], s = sc[2];
        for (int r = t; r <= iu2; r += blockDim.x) {  // rot on the right, rows 0..iu
          const double xv = AT(r, iu2 - 1), yv = AT(r, iu2);
          AT(r, iu2 - 1) = c * xv - s * yv;
          AT(r, iu2) = s * xv + c * yv;
        }
        __syncthreads();
        if (t == 0) AT(iu2, iu2 - 1) = 0;
      }
      if (t == 0 && iu2 > 1) AT(iu2 - 1, iu2 - 2) = 0;
      continue;  // the loop head synchronises
    }
    // ---- Francis QR step on the window [il, iuq]
    const int iuq = si[5];
    if (si[4]) {
      const double amt = sc[3];
      for (int i = t; i <= iuq; i += blockDim.x) AT(i, i) -= amt;
      __syncthreads();
    }
    // initFrancisQRStep: the largest im in (il, iuq-2] passing the
    // small-subdiagonal test, else il -- every candidate tested at once
    const double sh0 = sc[4], sh1 = sc[5], sh2 = sc[6];
    // (candidates m = iuq - 2 - t - 64k, k < 4: entries loaded at once, then
    // tested in descending order)
    cand = il;
    {
      double tmm[4] = {0, 0, 0, 0}, t10[4] = {0, 0, 0, 0}, t01[4] = {0, 0, 0, 0}, t11[4] = {0, 0, 0, 0};
      double t21[4] = {0, 0, 0, 0}, tm1[4] = {0, 0, 0, 0}, tpp[4] = {0, 0, 0, 0};
      if (n >= 4)  // otherwise there is no candidate m > il >= 0 with m + 2 <= iuq
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int m = min(max(iuq - 2 - t - 64 * k, 1), n - 3);  // in range; unused when m <= il
        tmm[k] = AT(m, m);
        t10[k] = AT(m + 1, m);
        t01[k] = AT(m, m + 1);
        t11[k] = AT(m + 1, m + 1);
        t21[k] = AT(m + 2, m + 1);
        tm1[k] = AT(m, m - 1);
        tpp[k] = AT(m - 1, m - 1);
      }
      bool found = false;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int m = iuq - 2 - t - 64 * k;
        if (m > il && !found) {
          const double Tmm = tmm[k];
          const double r = sh0 - Tmm, s = sh1 - Tmm;
          const double v0 = (r * s - sh2) / t10[k] + t01[k];
          const double v1 = t11[k] - Tmm - r - s;
          const double v2 = t21[k];
          const double lhs = tm1[k] * (fabs(v1) + fabs(v2));
          const double rhs = v0 * (fabs(tpp[k]) + fabs(Tmm) + fabs(t11[k]));
          if (fabs(lhs) < eps * rhs) { cand = m; found = true; }
        }
      }
    }
    const int im = block_reduce(cand, MaxOp(), ired);
    if (t == 0)
      for (int k = im; k <= iuq - 2; k++) work += 12.0 * ((iuq - k + 1) + ((iuq < k + 3 ? iuq : k + 3) + 1));
    double fv0, fv1, fv2;
    {
      const double Tmm = AT(im, im);
      const double r = sh0 - Tmm, s = sh1 - Tmm;
      fv0 = (r * s - sh2) / AT(im + 1, im) + AT(im, im + 1);
      fv1 = AT(im + 1, im + 1) - Tmm - r - s;
      fv2 = AT(im + 2, im + 1);
    }
    // Bulge chase in windows of QB steps.  Window steps [k0, k1) read and
    // update only entries inside R = [a, b) x [a, b) (a = k0-1, b = k1+3),
    // which live in LDS while wavefront 0 chases the bulge through them;
    // afterwards the window's reflectors are applied, in step order, to the
    // strips outside R (rows k0..k1+1 right of R: left reflectors only;
    // columns k0..k1+1 above R: right reflectors only).  Every entry
    // therefore receives the same reflectors in the same order as in the
    // unwindowed chase.
    for (int k0 = im; k0 <= iuq - 2; k0 += QB) {
      const int k1 = (k0 + QB < iuq - 1) ? k0 + QB : iuq - 1;
      const int a = k0 - 1 > 0 ? k0 - 1 : 0;
      const int b = (k1 + 3 < iuq + 1) ? k1 + 3 : iuq + 1;
      const int wd = b - a;
      {  // 32 x 32 tile, RS rows per block pass, all loads in flight
        int rb = t >> 5;
        asm volatile("" : "+v"(rb));  // per-row values are formed here, not hoisted out of the sweep loop
        const int cl = t & 31;
        const int cc = min(cl, wd - 1);  // out-of-window lanes re-read in-window entries (unused)
        double v[QW / RS];
#pragma unroll
        for (int it = 0; it < QW / RS; it++) v[it] = A[(unsigned)((a + min(rb + RS * it, wd - 1)) * n + a + cc)];
#pragma unroll
        for (int it = 0; it < QW / RS; it++) {
          const int r = rb + RS * it;
          if (r < wd && cl < wd) Rw[r * QWP + cl] = v[it];
        }
      }
      __syncthreads();
#define RW(r, c) Rw[((r) - a) * QWP + ((c) - a)]
      if (t < 64) {
        for (int k = k0; k < k1; ++k) {
          const bool first = (k == im);
          double x0, x1, x2;
          if (first) { x0 = fv0; x1 = fv1; x2 = fv2; }
          else { x0 = RW(k, k - 1); x1 = RW(k + 1, k - 1); x2 = RW(k + 2, k - 1); }
          // makeHouseholder of (x0, x1, x2), every lane
          double tail = 0;
          tail += x1 * x1;
          tail += x2 * x2;
          double tau = 0, e0 = 0, e1 = 0, beta = x0;
          if (tail != 0.0) {
            double bb = sqrt(x0 * x0 + tail);
            if (x0 >= 0) bb = -bb;
            e0 = x1 / (x0 - bb);
            e1 = x2 / (x0 - bb);
            tau = (bb - x0) / bb;
            beta = bb;
          }
          wave_sync();
          if (t == 0) {
            rf[k - k0][0] = beta != 0 ? tau : 0.0;
            rf[k - k0][1] = e0;
            rf[k - k0][2] = e1;
            rf[k - k0][3] = beta != 0 ? 1.0 : 0.0;
          }
          if (beta != 0) {
            if (t == 0) {
              if (first && k > il) RW(k, k - 1) = -RW(k, k - 1);
              else if (!first) RW(k, k - 1) = beta;
            }
            const int c = k + t;
            if (c < b) {
              double tt = e0 * RW(k + 1, c);
              tt += e1 * RW(k + 2, c);
              tt += RW(k, c);
              RW(k, c) -= tau * tt;
              RW(k + 1, c) -= (tau * e0) * tt;
              RW(k + 2, c) -= (tau * e1) * tt;
            }
            wave_sync();
            const int rlast = (iuq < k + 3 ? iuq : k + 3);
            const int r = a + t;
            if (r <= rlast) {
              double tt = RW(r, k + 1) * e0;
              tt += RW(r, k + 2) * e1;
              tt += RW(r, k);
              RW(r, k) -= tau * tt;
              RW(r, k + 1) -= (tau * tt) * e0;
              RW(r, k + 2) -= (tau * tt) * e1;
            }
            wave_sync();
          }
        }
      }
      __syncthreads();
      // the window's steps whose reflector is applied (beta != 0), one bit per
      // step: a uniform mask, so the strip loops branch on it without an LDS
      // round trip per step
      const unsigned long long fon = __ballot((t & 63) < QB && rf[min(t & 63, QB - 1)][3] != 0.0);
      for (int r = t >> 5; r < wd; r += RS)
        if ((t & 31) < wd) AT(a + r, a + (t & 31)) = Rw[r * QWP + (t & 31)];
#undef RW
      // strip right of R: rows k0 .. k1+1, columns b .. iuq (left reflectors).
      // Reflector q touches rows q..q+2 of the strip, so after it row q is
      // final: each thread streams down its column with a 3-value window.
      const int nstep = k1 - k0;
      for (int c = b + t; c <= iuq; c += blockDim.x) {
        gptr<double> p = A + ((long)k0 * n + c);
        double w[QB + 2];
        {
          gptr<const double> pl = p;  // running pointer: one live address
#pragma unroll
          for (int q = 0; q < QB + 2; q++) {
            w[q] = (q < nstep + 2) ? *pl : 0.0;
            pl += n;
          }
        }
        // every load has arrived before the first store: the stores below then
        // never stand in the way of a load's wait (the vector-memory counter
        // counts both)
        __builtin_amdgcn_s_waitcnt(0);
        gptr<double> ps = p;
        double u0 = w[0], u1 = w[1];
#pragma unroll
        for (int q = 0; q < QB; q++) {
          if (q < nstep) {
            double u2 = w[q + 2];
            if ((fon >> q) & 1) {
              const double tau = rf[q][0], e0 = rf[q][1], e1 = rf[q][2];
              double tt = e0 * u1;
              tt += e1 * u2;
              tt += u0;
              u0 -= tau * tt;
              u1 -= (tau * e0) * tt;
              u2 -= (tau * e1) * tt;
            }
            *ps = u0;
            ps += n;
            u0 = u1;
            u1 = u2;
          }
          __asm__ volatile("" ::: "memory");  // keep the reflector loads step by step
        }
        ps[0] = u0;
        ps[n] = u1;
      }
      // strip above R: rows 0 .. a-1, columns k0 .. k1+1 (right reflectors).
      // Staged through the (now free) window buffer in chunks of CH rows with
      // coalesced loads and stores; one lane streams each row through the
      // reflectors in LDS.
      __syncthreads();
      for (int r0 = 0; r0 < a; r0 += CH) {  // a chunk fills at most the window buffer
        const int nr = (a - r0 < CH) ? a - r0 : CH;
        const int nc = nstep + 2;
        {
          int rb = t >> 5;
          asm volatile("" : "+v"(rb));
          const int cl = t & 31;
          const int cc = min(cl, nc - 1);
          double v[CH / RS];
#pragma unroll
          for (int it = 0; it < CH / RS; it++) v[it] = A[(unsigned)((r0 + min(rb + RS * it, nr - 1)) * n + k0 + cc)];
#pragma unroll
          for (int it = 0; it < CH / RS; it++) {
            const int rr = rb + RS * it;
            if (rr < nr && cl < nc) Rw[rr * QWP + cl] = v[it];
          }
        }
        __syncthreads();
        if (t < nr) {
          // the row's entries first (LDS reads in flight together; entries
          // past the window's columns are read but left as they are)
          double *row = Rw + t * QWP;
          double x[QB + 2];
#pragma unroll
          for (int q = 0; q < QB + 2; q++) x[q] = row[q];
          double u0 = x[0], u1 = x[1];
#pragma unroll
          for (int q = 0; q < QB; q++) {
            if (q < nstep) {
              double u2 = x[q + 2];
              if ((fon >> q) & 1) {
                const double tau = rf[q][0], e0 = rf[q][1], e1 = rf[q][2];
                double tt = u1 * e0;
                tt += u2 * e1;
                tt += u0;
                u0 -= tau * tt;
                u1 -= (tau * tt) * e0;
                u2 -= (tau * tt) * e1;
              }
              row[q] = u0;
              u0 = u1;
              u1 = u2;
            }
            __asm__ volatile("" ::: "memory");  // the reflector reads step by step
          }
          row[nstep] = u0;
          row[nstep + 1] = u1;
        }
        __syncthreads();
        for (int rr = t >> 5; rr < nr; rr += RS)
          if ((t & 31) < nc) AT(r0 + rr, k0 + (t & 31)) = Rw[rr * QWP + (t & 31)];
        __syncthreads();
      }
    }
    if (t == 0) {
      double x[2] = {AT(iuq - 1, iuq - 2), AT(iuq, iuq - 2)};
      double e1[1], tau, beta;
      make_hh(x, 2, e1, &tau, &beta);
      si[6] = beta != 0;
      if (beta != 0) { AT(iuq - 1, iuq - 2) = beta; work += 6.0 * (2 + iuq + 1); }
      sc[3] = tau; sc[7] = e1[0];
    }
    __syncthreads();
    if (si[6]) {
      const double tau = sc[3], e0 = sc[7];
      if (t == 0)
        for (int c = iuq - 1; c <= iuq; c++) {
          double tt = e0 * AT(iuq, c);
          tt += AT(iuq - 1, c);
          AT(iuq - 1, c) -= tau * tt;
          AT(iuq, c) -= (tau * e0) * tt;
        }
      __syncthreads();
      for (int r = t; r <= iuq; r += blockDim.x) {
        double tt = AT(r, iuq) * e0;
        tt += AT(r, iuq - 1);
        AT(r, iuq - 1) -= tau * tt;
        AT(r, iuq) -= (tau * tt) * e0;
      }
    }
    for (int i = im + 2 + t; i <= iuq; i += blockDim.x) {
      AT(i, i - 2) = 0;
      if (i > im + 2) AT(i, i - 3) = 0;
    }
  }
  __syncthreads();
  // eigenvalue 2 in Schur-diagonal order (EigenSolver::compute, EigenSolver.h:364-381)
  if (t == 0) {
    double f = 0.0;
    if (n >= 3) {
      int i = 0;
      double re2 = 0, im2 = 0;
      while (i < n && i <= 2) {
        if (i == n - 1 || AT(i + 1, i) == 0) {
          if (i == 2) { re2 = AT(i, i); im2 = 0; }
          ++i;
        } else {
          const double p = 0.5 * (AT(i, i) - AT(i + 1, i + 1));
          const double z = sqrt(fabs(p * p + AT(i + 1, i) * AT(i, i + 1)));
          if (i == 2) { re2 = AT(i + 1, i + 1) + p; im2 = z; }
          if (i + 1 == 2) { re2 = AT(i + 1, i + 1) + p; im2 = -z; }
          i += 2;
        }
      }
      f = (im2 == 0) ? sqrt(re2) : 0.0;
    }
    H.feat[angle * 14 + 13] = f;
    H.work[angle] = work;
  }
#undef AT
}

// mean and range over the 4 angles, reference output order (haralick.cpp:96-143)
__global__ void k_har_final(const HarDesc *hd, OutRef o) {
  const HarDesc H = hd[blockIdx.x];
  if (threadIdx.x != 0) return;
  double temp[28];
  for (int a = 0; a < 14; a++) {
    double s = 0, mn = kInf, mx = -kInf;
    for (int g = 0; g < 4; g++) {
      const double v = H.feat[g * 14 + a];
      s += v;
      if (v < mn) mn = v;
      if (v > mx) mx = v;
    }
    temp[a] = s / 4;
    temp[a + 14] = mx - mn;
  }
  const int ord[28] = {0, 14, 1, 15, 2, 16, 9, 23, 10, 24, 8, 22, 11, 25,
                       4, 18, 13, 27, 12, 26, 5, 19, 7, 21, 6, 20, 3, 17};
  double *out = o.rows[blockIdx.x];
  for (int a = 0; a < 28; a++) out[a] = temp[ord[a]];
}

void launch_haralick_prep(const Img *imgs, const Stats *st, const HarDesc *hd, int nimg, int maxn, hipStream_t s) {
  int blocks = (maxn + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  k_har_quant<<<dim3(blocks, nimg), 256, 0, s>>>(imgs, st, hd);
  k_har_lut<<<nimg, 64, 0, s>>>(hd);
  static const bool lds_ok = hipFuncSetAttribute((const void *)k_har_cooc, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                  256 * 257 / 2 * sizeof(uint32_t)) == hipSuccess;
  (void)lds_ok;
  k_har_cooc<<<dim3(4, nimg), 1024, 256 * 257 / 2 * sizeof(uint32_t), s>>>(imgs, hd);
  k_har_stats<<<dim3(4, nimg), 256, 0, s>>>(hd);
  k_har_q<<<xcd_grid(QRB, 4 * nimg), 256, 0, s>>>(hd, nimg);
}
// Entries in decreasing order of tones (ng), for the eigen kernels: their
// cost grows with ng (about ng^2 for the Schur sweeps, ng^3 for the
// Hessenberg reduction), and ng ranges over 10..256 within a batch, so
// dispatching the largest matrices first keeps the last wave of blocks
// short (longest processing time first).  Every matrix is computed on its
// own, so the order changes no result.  Thread b owns key b = 256 - ng and
// scans the entries in index order (keys staged in LDS, chunk by chunk), so
// ties keep index order and the order is the same on every run.
constexpr int ORDER_CHUNK = 16384;
__global__ void __launch_bounds__(320) k_har_order(const int *ng, int nimg, int *order) {
  __shared__ __align__(16) uint16_t key[ORDER_CHUNK];
  __shared__ int cnt[257];
  const int t = threadIdx.x;
  for (int b = t; b < 257; b += blockDim.x) cnt[b] = 0;
  __syncthreads();
  for (int e = t; e < nimg; e += blockDim.x) atomicAdd(&cnt[256 - min(max(ng[e], 0), 256)], 1);
  __syncthreads();
  if (t == 0) {
    int sum = 0;
    for (int b = 0; b < 257; b++) {
      const int c = cnt[b];
      cnt[b] = sum;
      sum += c;
    }
  }
  __syncthreads();
  int pos = t < 257 ? cnt[t] : 0;
  for (int base = 0; base < nimg; base += ORDER_CHUNK) {
    const int m = min(ORDER_CHUNK, nimg - base), m8 = (m + 7) & ~7;
    __syncthreads();  // the previous chunk's keys are consumed
    for (int e = t; e < m8; e += blockDim.x)  // pad to eight keys with a key no thread owns
      key[e] = e < m ? (uint16_t)(256 - min(max(ng[base + e], 0), 256)) : (uint16_t)0xffff;
    __syncthreads();
    if (t < 257)
      for (int e = 0; e < m8; e += 8) {  // eight keys per LDS read
        const uint4 kk = *(const uint4 *)(key + e);
        const unsigned w[4] = {kk.x, kk.y, kk.z, kk.w};
#pragma unroll
        for (int j = 0; j < 8; j++)
          if ((int)((w[j >> 1] >> (16 * (j & 1))) & 0xffff) == t) order[pos++] = base + e + j;
      }
  }
}
void launch_haralick_order(const int *ng, int nimg, int *order, hipStream_t s) {
  k_har_order<<<1, 320, 0, s>>>(ng, nimg, order);
}
void launch_haralick_hess(const HarDesc *hd, const int *order, int nimg, hipStream_t s) {
  k_har_hess<<<dim3(4, nimg), 256, 0, s>>>(hd, order);
}
void launch_haralick_qr(const HarDesc *hd, const int *order, int nimg, hipStream_t s) {
  // up to 1024 matrices: one 256-thread block each (all resident at once);
  // beyond, one wavefront each
  if (4 * nimg <= 1024) k_har_qr<256><<<dim3(4, nimg), 256, 0, s>>>(hd, order);
  else k_har_qr<64><<<dim3(4, nimg), 64, 0, s>>>(hd, order);
}
void launch_haralick_final(const HarDesc *hd, OutRef o, int nimg, hipStream_t s) {
  k_har_final<<<nimg, 64, 0, s>>>(hd, o);
}

// ======================================================================= Tamura
// tamura.cpp:10-299.  scratch per image (5P doubles): [0, P) normalised
// plane, [P, 2P) summed-area table (laufendeSumme), [2P, 3P) phi,
// [3P, 4P) per-pixel sum_r terms, [4P, 4.5P) best coarseness scale (int).
// (x - min) / max per pixel; planes are 256-byte aligned (arena), so each
// thread streams two double pairs through the global address space
typedef double tam_d2 __attribute__((ext_vector_type(2)));
__global__ void __launch_bounds__(256) k_tam_norm(const Img *imgs, const Stats *st, double *const *scr) {
  const Img im = imgs[blockIdx.y];
  const Stats S = st[blockIdx.y];
  const long n = (long)im.h * im.w, n4 = n >> 2;
  const gptr<const tam_d2> src = (gptr<const tam_d2>)to_global((const double *)im.p);
  const gptr<tam_d2> dst = (gptr<tam_d2>)to_global(scr[blockIdx.y]);
  const double mn = S.mn, mx = S.mx;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n4; i += (long)gridDim.x * blockDim.x) {
    tam_d2 a = src[2 * i], b = src[2 * i + 1];
    a.x = (a.x - mn) / mx;
    a.y = (a.y - mn) / mx;
    b.x = (b.x - mn) / mx;
    b.y = (b.y - mn) / mx;
    dst[2 * i] = a;
    dst[2 * i + 1] = b;
  }
  if (blockIdx.x == 0 && threadIdx.x < (n & 3)) {
    const long i = (n4 << 2) + threadIdx.x;
    scr[blockIdx.y][i] = (im.p[i] - mn) / mx;
  }
}

// summed-area table with the reference's exact recurrence
//   S(y,x) = ((I(y,x) + S(y,x-1)) + S(y-1,x)) - S(y-1,x-1)
// evaluated as an anti-diagonal wavefront: one wave per image, lane l owns
// row y0+l of a 64-row strip, the value above comes from lane l-1.
__global__ void __launch_bounds__(64) k_tam_sat(const Img *imgs, double *const *scr) {
  extern __shared__ double rowbuf[];
  const Img im = imgs[blockIdx.x];
  const int H = im.h, W = im.w;
  const long P = (long)H * W;
  // global, not FLAT: the LDS waits of each step leave the prefetched
  // inputs in flight
  const gptr<const double> I = to_global((const double *)scr[blockIdx.x]);
  const gptr<double> S = to_global(scr[blockIdx.x] + P);
  const int l = threadIdx.x;
  constexpr int PF = 16;  // inputs loaded ahead per lane
  for (int y0 = 0; y0 < H; y0 += 64) {
    const int y = y0 + l;
    const gptr<const double> Irow = I + (long)y * W;
    double cur = 0, up_prev = 0;
    double nx[PF];
#pragma unroll
    for (int k = 0; k < PF; k++) {
      const int x = k - l;
      nx[k] = (y < H && x >= 0 && x < W) ? Irow[x] : 0.0;
    }
    for (int t0 = 0; t0 < W + 63; t0 += PF) {
      double in[PF];
#pragma unroll
      for (int k = 0; k < PF; k++) in[k] = nx[k];
#pragma unroll
      for (int k = 0; k < PF; k++) {  // the next group's inputs, in flight during this group
        const int x = t0 + PF + k - l;
        nx[k] = (y < H && x >= 0 && x < W) ? Irow[x] : 0.0;
      }
#pragma unroll
      for (int k = 0; k < PF; k++) {
        const int t = t0 + k;
        if (t >= W + 63) break;
        const int x = t - l;
        const double from_left_lane = __shfl_up(cur, 1, 64);
        double up;
        if (l == 0) up = (y0 == 0 || x < 0 || x >= W) ? 0.0 : rowbuf[x];
        else up = (y < 1) ? 0.0 : from_left_lane;
        if (x >= 0 && x < W && y < H) {
          const double links = x < 1 ? 0.0 : cur;
          const double oben = y < 1 ? 0.0 : up;
          const double obenlinks = (y < 1 || x < 1) ? 0.0 : up_prev;
          const double v = in[k] + links + oben - obenlinks;
          S[(long)y * W + x] = v;
          cur = v;
          up_prev = up;
        }
        __syncthreads();
        if (l == 63 && x >= 0 && x < W) rowbuf[x] = cur;
        __syncthreads();
      }
    }
  }
}

// The window's pixel count is a power of two 2^e away from the borders;
// there the quotient is formed as a product with 2^-e, which is exact, so
// both are the correctly rounded value of the same real number.
__device__ __forceinline__ double local_mean(int x, int y, int k, gptr<const double> S, int dimx, int dimy) {
  const int k2 = k / 2;
  int starty = y - k2, startx = x - k2, stopy = y + k2 - 1, stopx = x + k2 - 1;
  if (starty < 0) starty = 0;
  if (startx < 0) startx = 0;
  if (stopx > dimx - 1) stopx = dimx - 1;
  if (stopy > dimy - 1) stopy = dimy - 1;
  const long r1 = (long)stopy * dimx, r0 = r1 - dimx;
  const double links = (startx - 1 < 0) ? 0 : S[r1 + startx - 1];
  const double oben = (starty - 1 < 0) ? 0 : S[r0 + startx];
  const double obenlinks = ((starty - 1 < 0) || (startx - 1 < 0)) ? 0 : S[r0 + startx - 1];
  const double unten = S[r1 + startx];
  const int counter = (stopy - starty + 1) * (stopx - startx + 1);
  const double num = unten - links - oben + obenlinks;
  if ((counter & (counter - 1)) == 0) {
    const int e = __builtin_ctz((unsigned)counter);
    return num * __longlong_as_double((long long)(1023 - e) << 52);  // 2^-e
  }
  return num / counter;
}

// coarseness per pixel (best scale), directionality phi and the sums
struct TamAcc {
  double sum_r, m2, m4, mean;
  long long ksum;
  int kmin, kmax;
};

__global__ void __launch_bounds__(256) k_tam_pix(const Img *imgs, double *const *scr, int nimg, int per_img) {
  int ent, blk;
  if (!xcd_item(per_img, nimg, ent, blk)) return;  // one plane's SAT stays in one L2
  const Img im = imgs[ent];
  const int H = im.h, W = im.w;
  const long P = (long)H * W;
  const long i = blk * (long)blockDim.x + threadIdx.x;
  if (i >= P) return;
  const gptr<double> N = to_global(scr[ent]);  // global, not FLAT: vector-memory waits only
  const gptr<const double> S = (gptr<const double>)(N + P);
  const gptr<double> phi = N + 2 * P;
  const int y = i / W, x = i - (long)y * W;
  // coarseness: argmax over scales of E_h / E_v (tamura.cpp:183-238)
  double maxE = 0;
  int maxk = 0;
  int len = 1;
  for (int k = 1; k <= 7; ++k) {
    const int k2 = len;
    len *= 2;
    double eh = 0, ev = 0;
    if (x + k2 < W && x - k2 >= 0)
      eh = fabs(local_mean(x + k2, y, len, S, W, H) - local_mean(x - k2, y, len, S, W, H));
    if (y + k2 < H && y - k2 >= 0)
      ev = fabs(local_mean(x, y + k2, len, S, W, H) - local_mean(x, y - k2, len, S, W, H));
    if (eh > maxE) { maxE = eh; maxk = k; }
    if (ev > maxE) { maxE = ev; maxk = k; }
  }
  ((gptr<int>)(N + 4 * P))[i] = maxk;
  // directionality filters after the reference's assignments (tamura.cpp:50-57):
  // H = [-1 2 1; -2 0 2; -1 -2 -1], V = [1 0 0; 0 0 0; -1 0 0], correlation order
  // (column offset outer, row offset inner), zero weights skipped (exact).
  auto px = [&](int yy, int xx) -> double { return N[(long)yy * W + xx]; };
  double dh = 0.0, dv = 0.0;
  const bool xm = x - 1 >= 0, xp = x + 1 < W, ym = y - 1 >= 0, yp = y + 1 < H;
  if (xm) {
    if (ym) dh += -1 * px(y - 1, x - 1);
    dh += -2 * px(y, x - 1);
    if (yp) dh += -1 * px(y + 1, x - 1);
  }
  if (ym) dh += 2 * px(y - 1, x);
  if (yp) dh += -2 * px(y + 1, x);
  if (xp) {
    if (ym) dh += 1 * px(y - 1, x + 1);
    dh += 2 * px(y, x + 1);
    if (yp) dh += -1 * px(y + 1, x + 1);
  }
  if (xm) {
    if (ym) dv += 1 * px(y - 1, x - 1);
    if (yp) dv += -1 * px(y + 1, x - 1);
  }
  double ph = 0.0;
  if (dh >= 0.0001) ph = atan(dv / dh) + (M_PI / 2.0 + 0.001);
  phi[i] = ph;
  N[3 * P + i] = (dh >= 0.0001) ? (dh * dh + dv * dv) + ph * ph : 0.0;
}

__global__ void __launch_bounds__(1024) k_tam_final(const Img *imgs, double *const *scr, OutRef o) {
  __shared__ double shd[16];
  __shared__ long long shl[16];
  __shared__ int hk[3];
  __shared__ int hd[125];
  const Img im = imgs[blockIdx.x];
  const int H = im.h, W = im.w;
  const long P = (long)H * W;
  const double *N = scr[blockIdx.x];
  const double *phi = N + 2 * P;
  const double *srt = N + 3 * P;
  const int *kb = (const int *)(N + 4 * P);
  const int t = threadIdx.x;
  // contrast: Moments4 of the normalised plane (tamura.cpp:10-28)
  double s = 0;
  for (long i = t; i < P; i += blockDim.x) s += N[i];
  s = block_reduce(s, SumOp(), shd);
  const double mean = s / P;
  double m2 = 0, m4 = 0;
  for (long i = t; i < P; i += blockDim.x) { const double d = N[i] - mean, d2 = d * d; m2 += d2; m4 += d2 * d2; }
  m2 = block_reduce(m2, SumOp(), shd);
  m4 = block_reduce(m4, SumOp(), shd);
  // coarseness sums and min/max of best scale
  long long ks = 0;
  int kmn = 99, kmx = -1;
  double pmn = 1e308, pmx = -1e308, sr = 0;
  for (long i = t; i < P; i += blockDim.x) {
    const int k = kb[i];
    ks += k;
    kmn = min(kmn, k);
    kmx = max(kmx, k);
    pmn = fmin(pmn, phi[i]);
    pmx = fmax(pmx, phi[i]);
    sr += srt[i];
  }
  ks = block_reduce(ks, SumOp(), shl);
  kmn = (int)block_reduce((long long)kmn, MinOp(), shl);
  kmx = (int)block_reduce((long long)kmx, MaxOp(), shl);
  pmn = block_reduce(pmn, MinOp(), shd);
  pmx = block_reduce(pmx, MaxOp(), shd);
  sr = block_reduce(sr, SumOp(), shd);
  if (t < 3) hk[t] = 0;
  if (t < 125) hd[t] = 0;
  __syncthreads();
  const double ksc = hist_scale((double)kmn, (double)kmx, 3), psc = hist_scale(pmn, pmx, 125);
  // the 3-bin coarseness histogram by ballot per wave (counts in scalar
  // registers, one LDS add per bin and wave); the 125-bin one by LDS atomics
  int wk0 = 0, wk1 = 0, wk2 = 0;
  for (long i0 = 0; i0 < P; i0 += blockDim.x) {
    const long i = i0 + t;
    const bool ok = i < P;
    int b = -1;
    if (ok) {
      b = hist_bin((double)kb[i], (double)kmn, ksc, 3);
      atomicAdd(&hd[hist_bin(phi[i], pmn, psc, 125)], 1);
    }
    wk0 += __popcll(__ballot(b == 0));
    wk1 += __popcll(__ballot(b == 1));
    wk2 += __popcll(__ballot(b == 2));
  }
  if ((t & 63) == 0) {
    atomicAdd(&hk[0], wk0);
    atomicAdd(&hk[1], wk1);
    atomicAdd(&hk[2], wk2);
  }
  __syncthreads();
  if (t == 0) {
    double *out = o.rows[blockIdx.x];
    // coarseness (tamura.cpp:240-262); the divisor is unsigned arithmetic
    const unsigned xD = W, yD = H;
    double sum = (double)ks;
    if (yD == 32 || xD == 32) sum /= ((xD + 1 - 32) * (yD + 1 - 32));
    else sum /= ((yD - 32) * (xD - 32));
    double hm = 0;
    for (int k = 0; k < 3; k++) if (hk[k] > hm) hm = hk[k];
    // directionality (tamura.cpp:75-101)
    double mxh = 0.0;
    long fmx = 0;
    for (long a = 0; a < 125; a++) if (hd[a] > mxh) { mxh = hd[a]; fmx = a; }
    double ssum = 0;
    for (long a = 0; a < 125; a++) ssum += hd[a] * (double)((a + 1 - fmx) * (a + 1 - fmx));
    const double dir = fabs(log(ssum / sr + 0.0000001));
    // contrast
    double con;
    const double stdv = P > 2 ? sqrt(m2 / (P - 1)) : 0.0;
    const double kurt = P > 4 ? (P * m4) / (m2 * m2) : 0.0;
    if (stdv < kDblEps) con = 0;
    else con = stdv / pow(kurt / pow(stdv, 4), 0.25);
    out[0] = hk[0] / hm;
    out[1] = hk[1] / hm;
    out[2] = hk[2] / hm;
    out[3] = con;
    out[4] = dir;
    out[5] = sum;
  }
}

void launch_tamura(const Img *imgs, const Stats *st, double *const *scratch, OutRef o, int nimg, int maxh, int maxw,
                   hipStream_t s) {
  const long maxn = (long)maxh * maxw;
  int blocks = (int)((maxn + 255) / 256);
  const int nblocks = (int)((maxn / 4 + 255) / 256);
  k_tam_norm<<<dim3(nblocks < 1 ? 1 : (nblocks > 1024 ? 1024 : nblocks), nimg), 256, 0, s>>>(imgs, st, scratch);
  k_tam_sat<<<nimg, 64, maxw * sizeof(double), s>>>(imgs, scratch);
  k_tam_pix<<<xcd_grid(blocks, nimg), 256, 0, s>>>(imgs, scratch, nimg, blocks);
  k_tam_final<<<nimg, 1024, 0, s>>>(imgs, scratch, o);
}

// ================================================================= Comb moments
// CombFirst4Moments2D (CombFirst4Moments.cpp:78-217): 4 combs x 20 lines, one
// thread per line running the reference's sequential Moments4 over the
// line's pixels in raster order.  iis[img*20 + j] are the 20 rounded line
// offsets (host-computed with the reference's FP accumulation), NaN = unused.
__global__ void __launch_bounds__(128) k_comb(const Img *imgs, const double *iis, OutRef o) {
  __shared__ double mm[4][4][20];
  const Img im = imgs[blockIdx.x];
  const long m = im.h, n = im.w;
  const long n2 = n / 2, m2 = m / 2;
  const int t = threadIdx.x;
  if (t < 80) {
    const int comb = t / 20, li = t % 20;
    const double iid = iis[blockIdx.x * 20 + li];
    double z[4] = {0, 0, 0, 0};
    if (iid == iid) {
      const long ii = (long)iid;
      double mean = 0, M2 = 0, M3 = 0, M4 = 0;
      size_t cnt = 0;
      auto add = [&](double x) {
        if (!moments_ok(x)) return;
        const size_t n1 = cnt;
        cnt = cnt + 1;
        const double delta = x - mean;
        const double delta_n = delta / cnt;
        const double delta_n2 = delta_n * delta_n;
        const double term1 = delta * delta_n * n1;
        mean = mean + delta_n;
        M4 = M4 + term1 * delta_n2 * (cnt * cnt - 3 * cnt + 3) + 6 * delta_n2 * M2 - 4 * delta_n * M3;
        M3 = M3 + term1 * delta_n * (cnt - 2) - 3 * delta_n * M2;
        M2 = M2 + term1;
      };
      if (comb == 0) {         // |I + ii - J| < 1  ->  x = y + ii
        for (long y = 0; y < m; y++) { const long x = y + ii; if (x >= 0 && x < n) add(im.p[y * n + x]); }
      } else if (comb == 1) {  // J1 = n - x  ->  x = n - 1 - y - ii
        for (long y = 0; y < m; y++) { const long x = n - 1 - y - ii; if (x >= 0 && x < n) add(im.p[y * n + x]); }
      } else if (comb == 2) {  // vertical: x = n2 - 1 - ii
        const long x = n2 - 1 - ii;
        if (x >= 0 && x < n) for (long y = 0; y < m; y++) add(im.p[y * n + x]);
      } else {                 // horizontal: y = m2 - 1 - ii
        const long y = m2 - 1 - ii;
        if (y >= 0 && y < m) for (long x = 0; x < n; x++) add(im.p[y * n + x]);
      }
      z[0] = mean;
      z[1] = cnt > 2 ? sqrt(M2 / (cnt - 1)) : 0.0;
      z[2] = cnt > 3 ? (sqrt((double)cnt) * M3) / pow(M2, 1.5) : 0.0;
      z[3] = cnt > 4 ? (cnt * M4) / (M2 * M2) : 0.0;
    }
    for (int a = 0; a < 4; a++) mm[comb][a][li] = z[a];
  }
  __syncthreads();
  if (t == 0) {
    double vec[48];
    int vi = 0;
    for (int comb = 0; comb < 4; comb++)
      for (int a = 0; a < 4; a++) {
        double hmin = kInf, hmax = -kInf;
        for (int b = 0; b < 20; b++) { const double v = mm[comb][a][b]; if (v > hmax) hmax = v; if (v < hmin) hmin = v; }
        const double hs = hist_scale(hmin, hmax, 3);
        double bins[3] = {0, 0, 0};
        for (int b = 0; b < 20; b++) bins[hist_bin(mm[comb][a][b], hmin, hs, 3)] += 1.0;
        for (int b = 0; b < 3; b++) vec[vi++] = bins[b];
      }
    const int ord[48] = {45, 46, 47, 36, 37, 38, 42, 43, 44, 39, 40, 41, 33, 34, 35, 24,
                         25, 26, 30, 31, 32, 27, 28, 29, 9, 10, 11, 0, 1, 2, 6, 7,
                         8, 3, 4, 5, 21, 22, 23, 12, 13, 14, 18, 19, 20, 15, 16, 17};
    double *out = o.rows[blockIdx.x];
    for (int a = 0; a < 48; a++) out[a] = vec[ord[a]];
  }
}

void launch_comb(const Img *imgs, const double *iis, OutRef o, int nimg, hipStream_t s) {
  k_comb<<<nimg, 128, 0, s>>>(imgs, iis, o);
}

// ======================================================================== Radon
// radon.cpp:14-91 restated as a gather: each projection bin r of angle a sums,
// in the reference's (n, m) scan order, pixelLow of sub-pixels with
// floor(rIdx) == r and (pixel - pixelLow) of those with floor(rIdx) == r-1.
// floor(rIdx(n, m)) is non-increasing in m, so each column n contributes a
// contiguous m-range found by binary search.  trig[4][2] = {cos, sin} (glibc).
__device__ __forceinline__ double radon_ridx(int n, int m, int M, int N, double cs, double sn, int rFirst) {
  const int xO = (N - 1) / 2, yO = (M - 1) / 2;
  const double xT = (-xO - 0.25) + 0.5 * n;
  const double yT = (-yO - 0.25) + 0.5 * (2 * M - 1 - m);
  const double xc = xT * cs - rFirst;
  const double ys = yT * sn;
  return xc + ys;
}

// Block = (angle, group of 256 bins, image); thread = bin r.  The y*sin
// table of the reference (radon.cpp:58-63) is built in LDS and x*cos
// recomputed with the reference's operations, so rIdx is bit-identical.
// floor(rIdx) is monotone in the sub-column n (x*cos is) and non-increasing
// in the sub-row m, so the sub-columns touching bin r form an interval
// (binary search) and, per sub-column, the sub-rows with floor(rIdx) in
// {r-1, r} form a range that moves slowly with n (tracked by stepping).
// Each bin's terms are added in the reference's order: sub-columns
// ascending, sub-rows ascending within one.
//   At angle 0 a bin is a pixel column and the lanes of a wave read
// adjacent pixels of one row, so pixels are read from global memory
// directly.  At the other angles a lane follows its own slanted line, so a
// direct read touches one cache line per lane; there the block streams the
// plane through LDS in chunks of whole pixel columns (rows limited to the
// band its bins can reach, with a direct read as the fallback outside it),
// and every lane takes its sub-columns of a chunk in order.
// doubles of the LDS chunk buffer: 16 KB keeps 8 blocks per CU; the
// kernel is bound by each bin's sequential sum, and larger chunks (lower
// occupancy) measured slower
constexpr int RADON_TILE = 2048;
__device__ __forceinline__ double radon_xc(int nn, int xO, double cs, int rFirst) {
  return ((-xO - 0.25) + 0.5 * nn) * cs - rFirst;
}

__global__ void __launch_bounds__(256) k_radon_proj(const Img *imgs, const double *trig, double *const *proj,
                                                    int nbg, int nimg, int ysz, int a0, int na) {
  extern __shared__ double lds[];
  double *yS = lds, *tile = lds + ysz;
  __shared__ int ired[8];
  int ent, sub;
  if (!xcd_item(na * nbg, nimg, ent, sub)) return;
  const Img im = imgs[ent];
  const int M = im.h, N = im.w;
  const int rLast = (int)ceil(sqrt((double)((N - 1 - (N - 1) / 2) * (N - 1 - (N - 1) / 2)) +
                                   (double)((M - 1 - (M - 1) / 2) * (M - 1 - (M - 1) / 2)))) + 1;
  const int rFirst = -rLast, rSize = rLast - rFirst + 1;
  const int a = a0 + sub / nbg, r0 = (sub % nbg) * 256;
  if (r0 >= rSize) return;
  const double cs = trig[2 * a], sn = trig[2 * a + 1];
  const int xO = (N - 1) / 2, yO = (M - 1) / 2;
  const int M2 = 2 * M, N2 = 2 * N;
  for (int p = threadIdx.x; p < M2; p += blockDim.x) yS[p] = ((-yO - 0.25) + 0.5 * (M2 - 1 - p)) * sn;
  __syncthreads();
  const int r = r0 + threadIdx.x;
  const bool act = r < rSize;
  const double yS0 = yS[0], ySL = yS[M2 - 1];
  // column nn is skipped when floor(xc + yS0) < r-1 (A) or floor(xc + ySL) > r (B)
  auto predA = [&](int nn) { return (int)(radon_xc(nn, xO, cs, rFirst) + yS0) < r - 1; };
  auto predB = [&](int nn) { return (int)(radon_xc(nn, xO, cs, rFirst) + ySL) > r; };
  int lo = 0, hi = N2;
  if (!act) {
    hi = 0;
  } else if (cs > 0) {  // A: true then false; B: false then true
    int l = 0, h = N2;
    while (l < h) { const int mid = (l + h) >> 1; if (!predA(mid)) h = mid; else l = mid + 1; }
    lo = l;
    l = lo; h = N2;
    while (l < h) { const int mid = (l + h) >> 1; if (predB(mid)) h = mid; else l = mid + 1; }
    hi = l;
  } else if (cs < 0) {  // B: true then false; A: false then true
    int l = 0, h = N2;
    while (l < h) { const int mid = (l + h) >> 1; if (!predB(mid)) h = mid; else l = mid + 1; }
    lo = l;
    l = lo; h = N2;
    while (l < h) { const int mid = (l + h) >> 1; if (predA(mid)) h = mid; else l = mid + 1; }
    hi = l;
  } else if (predA(0) || predB(0)) {
    hi = 0;
  }
  const gptr<const double> P = to_global((const double *)im.p);
  double acc = 0.0;
  int m0 = 0, m1 = 0;
  // sub-rows [m0, m1) of column nn whose bin is r or r-1; [m0, m1) moves
  // monotonically with nn
  auto range = [&](int nn, double xc) {
    if (nn == lo) {
      int l = 0, h = M2;
      while (l < h) { const int mid = (l + h) >> 1; if ((int)(xc + yS[mid]) <= r) h = mid; else l = mid + 1; }
      m0 = l;
      l = m0; h = M2;
      while (l < h) { const int mid = (l + h) >> 1; if ((int)(xc + yS[mid]) < r - 1) h = mid; else l = mid + 1; }
      m1 = l;
    } else {
      while (m0 > 0 && (int)(xc + yS[m0 - 1]) <= r) m0--;
      while (m0 < M2 && (int)(xc + yS[m0]) > r) m0++;
      if (m1 < m0) m1 = m0;
      while (m1 > m0 && (int)(xc + yS[m1 - 1]) < r - 1) m1--;
      while (m1 < M2 && (int)(xc + yS[m1]) >= r - 1) m1++;
    }
  };
  auto term = [&](double pixel, double xc, int mm) {
    if (pixel) {
      pixel *= 0.25;
      const double rIdx = (xc + yS[mm]);
      const int rLow = (int)rIdx;
      const double pixelLow = pixel * (1 - rIdx + rLow);
      if (rLow == r) acc += pixelLow;
      else acc += pixel - pixelLow;
    }
  };
  if (fabs(sn) < 0.5) {  // angle 0: lanes read adjacent pixels of a row
    for (int nn = lo; nn < hi; nn++) {
      const double xc = radon_xc(nn, xO, cs, rFirst);
      range(nn, xc);
      const gptr<const double> col = P + (nn >> 1);
#pragma unroll 4
      for (int mm = m0; mm < m1; mm++) term(col[(unsigned)((mm >> 1) * N)], xc, mm);
    }
    if (act) proj[ent][(long)a * rSize + r] = acc;
    return;
  }
  // pixel columns any bin of the block touches
  const int xlo = block_reduce(lo < hi ? lo >> 1 : N, MinOp(), ired);
  const int xhi = block_reduce(lo < hi ? (hi + 1) >> 1 : 0, MaxOp(), ired);
  const int Wc = max(1, min(32, RADON_TILE / (M + 1)));  // wc * ld <= RADON_TILE (ld <= M + 1)
  const int r1 = min(r0 + 256, rSize);
  int nn = lo;
  for (int x0 = xlo; x0 < xhi; x0 += Wc) {
    const int x1 = min(x0 + Wc, N), wc = x1 - x0;
    // rows the block's bins [r0, r1) can reach in sub-columns [2 x0, 2 x1):
    // floor(rIdx) in [r0 - 1, r1 - 1] with rIdx = xc + yS(mm) and
    // yS(mm) = (-yO - 0.25 + 0.5 (M2 - 1 - mm)) sn, sn > 0; two rows of margin
    const double xa = radon_xc(2 * x0, xO, cs, rFirst), xb = radon_xc(2 * x1 - 1, xO, cs, rFirst);
    const double xcmin = fmin(xa, xb), xcmax = fmax(xa, xb);
    const double c0 = -yO - 0.25;
    const double mlo = (M2 - 1) - 2 * ((r1 - xcmin) / sn - c0);
    const double mhi = (M2 - 1) - 2 * ((r0 - 1 - xcmax) / sn - c0);
    const int rb0 = max(0, (int)floor(fmax(mlo, -4.0) * 0.5) - 2);
    const int rb1 = min(M, (int)floor(fmin(mhi, (double)M2 + 4) * 0.5) + 3);
    const int band = max(0, rb1 - rb0);
    const int ld = band | 1;  // odd stride: column reads of adjacent rows fall in different banks
    __syncthreads();  // the previous chunk is consumed
    for (int e = threadIdx.x; e < wc * band; e += blockDim.x) {
      const int row = e / wc, c = e - row * wc;
      tile[c * ld + row] = P[(long)(rb0 + row) * N + x0 + c];
    }
    __syncthreads();
    const int nend = min(hi, 2 * x1);
    for (; nn < nend; nn++) {
      const double xc = radon_xc(nn, xO, cs, rFirst);
      range(nn, xc);
      const int c = (nn >> 1) - x0;
      for (int mm = m0; mm < m1; mm++) {
        const int row = mm >> 1;
        const double pixel = (row >= rb0 && row < rb1) ? tile[c * ld + row - rb0] : P[(long)row * N + (nn >> 1)];
        term(pixel, xc, mm);
      }
    }
  }
  if (act) proj[ent][(long)a * rSize + r] = acc;
}

// yTable / xTable entries are exact multiples of 0.25, so the closed forms
// above equal the reference's running sums.
__global__ void k_radon_hist(const Img *imgs, double *const *proj, OutRef o) {
  const Img im = imgs[blockIdx.x];
  const int M = im.h, N = im.w;
  const int rLast = (int)ceil(sqrt((double)((N - 1 - (N - 1) / 2) * (N - 1 - (N - 1) / 2)) +
                                   (double)((M - 1 - (M - 1) / 2) * (M - 1 - (M - 1) / 2)))) + 1;
  const int rSize = 2 * rLast + 1;
  if (threadIdx.x != 0) return;
  double v[12];
  int vi = 0;
  for (int a = 0; a < 4; a++) {
    const double *p = proj[blockIdx.x] + (long)a * rSize;
    double hmin = kInf, hmax = -kInf;
    for (int i = 0; i < rSize; i++) { if (p[i] > hmax) hmax = p[i]; if (p[i] < hmin) hmin = p[i]; }
    const double hs = hist_scale(hmin, hmax, 3);
    double bins[3] = {0, 0, 0};
    for (int i = 0; i < rSize; i++) bins[hist_bin(p[i], hmin, hs, 3)] += 1.0;
    for (int b = 0; b < 3; b++) v[vi++] = bins[b];
  }
  const int ord[12] = {0, 1, 2, 9, 10, 11, 3, 4, 5, 6, 7, 8};
  double *out = o.rows[blockIdx.x];
  for (int a = 0; a < 12; a++) out[a] = v[ord[a]];
}

// --------------------------------------------------- certified (order-free) form
// A projection bin of the reference is the sequential sum of its terms.  Any
// other summation of the same terms lies within gamma_K * sum|terms| of their
// exact sum, as does the reference's (K bounds the depth of both summation
// trees), so the reference's bin lies within 2 gamma_K sum|terms| of an
// order-free sum.  The 12 features only need each bin's 3-bin histogram slot
// and the extreme bins of each angle: k_radon_sum forms the order-free sums
// with a bound per bin, and k_radon_cert settles the extremes and every slot
// from them, recomputing in the reference's order the bins a bound leaves
// open.  The counts equal the reference's bit for bit.

// LDS of one wave's recomputation: the terms of one window, and per
// sub-column of a chunk the offset of its first term and its first sub-row
constexpr int RADON_RTB = 512;  // terms per window (8 per lane)
constexpr int RADON_RCH = 256;  // sub-columns per chunk (4 per lane)
struct RadonWaveLds {
  double tb[RADON_RTB];
  int off[RADON_RCH];
  int m0[RADON_RCH];
};

// Bin r of the projection at (cs, sn) in the reference's order (radon.cpp:
// 66-81: sub-columns ascending, sub-rows ascending within one), by the 64
// lanes of a wave (r uniform).  Per chunk of 256 sub-columns each lane finds
// the sub-rows of four (those whose floor(rIdx) is r or r-1, a range, as
// floor(rIdx) is non-increasing in the sub-row), a wave scan turns the counts
// into offsets, the lanes form the terms 512 at a time into LDS in the
// reference's order with its operations (8 pixel loads in flight per lane),
// and lane 0 adds them in that order.  A zero pixel, which the reference
// skips, gives a +0 term (the running sum is never -0, so adding +0 leaves it
// unchanged).  Every lane returns the sum.
__device__ double radon_bin_wave(gptr<const double> P, int M, int N, double cs, double sn, int rFirst, int r,
                                 RadonWaveLds &W) {
  const int lane = threadIdx.x & 63;
  const int xO = (N - 1) / 2, yO = (M - 1) / 2, M2 = 2 * M, N2 = 2 * N;
  auto yS = [&](int mm) { return ((-yO - 0.25) + 0.5 * (M2 - 1 - mm)) * sn; };
  const double yS0 = yS(0), ySL = yS(M2 - 1);
  auto predA = [&](int nn) { return (int)(radon_xc(nn, xO, cs, rFirst) + yS0) < r - 1; };
  auto predB = [&](int nn) { return (int)(radon_xc(nn, xO, cs, rFirst) + ySL) > r; };
  int lo = 0, hi = N2;
  if (cs > 0) {
    int l = 0, h = N2;
    while (l < h) { const int mid = (l + h) >> 1; if (!predA(mid)) h = mid; else l = mid + 1; }
    lo = l;
    l = lo; h = N2;
    while (l < h) { const int mid = (l + h) >> 1; if (predB(mid)) h = mid; else l = mid + 1; }
    hi = l;
  } else if (cs < 0) {
    int l = 0, h = N2;
    while (l < h) { const int mid = (l + h) >> 1; if (!predB(mid)) h = mid; else l = mid + 1; }
    lo = l;
    l = lo; h = N2;
    while (l < h) { const int mid = (l + h) >> 1; if (predA(mid)) h = mid; else l = mid + 1; }
    hi = l;
  } else if (predA(0) || predB(0)) {
    hi = 0;
  }
  double acc = 0.0;
  for (int cb = lo; cb < hi; cb += RADON_RCH) {
    int cnt[4], mlo[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int nn = cb + 4 * lane + k;
      int m0 = 0, m1 = 0;
      if (nn < hi) {
        const double xc = radon_xc(nn, xO, cs, rFirst);
        int l = 0, h = M2;
        while (l < h) { const int mid = (l + h) >> 1; if ((int)(xc + yS(mid)) <= r) h = mid; else l = mid + 1; }
        m0 = l;
        h = M2;
        while (l < h) { const int mid = (l + h) >> 1; if ((int)(xc + yS(mid)) < r - 1) h = mid; else l = mid + 1; }
        m1 = l;
      }
      cnt[k] = m1 - m0;
      mlo[k] = m0;
    }
    const int own = cnt[0] + cnt[1] + cnt[2] + cnt[3];
    int inc = own;
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(inc, o, 64);
      if (lane >= o) inc += t;
    }
    const int total = __shfl(inc, 63, 64);
    int ofs = inc - own;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      W.off[4 * lane + k] = ofs;
      W.m0[4 * lane + k] = mlo[k];
      ofs += cnt[k];
    }
    wave_sync();
    for (int w0 = 0; w0 < total; w0 += RADON_RTB) {
      const int wn = min(RADON_RTB, total - w0);
      int nj[8], mj[8];
      double px[8];
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const int g = w0 + min(lane + 64 * i, wn - 1);
        int j = 0;  // the sub-column holding term g: the last j with off[j] <= g
        for (int st = RADON_RCH / 2; st > 0; st >>= 1)
          if (W.off[j + st] <= g) j += st;
        nj[i] = cb + j;
        mj[i] = W.m0[j] + (g - W.off[j]);
        px[i] = P[(long)(mj[i] >> 1) * N + (nj[i] >> 1)];
      }
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const int t = lane + 64 * i;
        if (t < wn) {
          double pixel = px[i], term = 0.0;
          if (pixel) {
            pixel *= 0.25;
            const double rIdx = (radon_xc(nj[i], xO, cs, rFirst) + yS(mj[i]));
            const int rLow = (int)rIdx;
            const double pixelLow = pixel * (1 - rIdx + rLow);
            term = (rLow == r) ? pixelLow : pixel - pixelLow;
          }
          W.tb[t] = term;
        }
      }
      wave_sync();
      if (lane == 0) {
        int i = 0;
        for (; i + 8 <= wn; i += 8) {
          double t8[8];
#pragma unroll
          for (int k = 0; k < 8; k++) t8[k] = W.tb[i + k];
#pragma unroll
          for (int k = 0; k < 8; k++) acc += t8[k];
        }
        for (; i < wn; i++) acc += W.tb[i];
      }
      wave_sync();  // the window is consumed
    }
  }
  return __shfl(acc, 0, 64);
}

template <typename T, typename Op>
__device__ __forceinline__ T wave_all_reduce(T v, Op op) {
  for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o, 64));
  return v;
}

// Order-free projection sums.  Block = (group of NA angles, image plane); a
// thread walks pixel columns top to bottom (the lanes of a wave read adjacent
// pixels of a row).  Each sub-pixel's two terms are the reference's
// (radon.cpp:66-81: the same rIdx, rLow, pixelLow and pixel - pixelLow).  Down
// a column rIdx is non-increasing, and the four sub-pixels of a pixel span
// less than one unit of it, so a pixel's terms fall in bins pb..pb+2 (pb =
// its smallest rLow) and pb never rises: the thread keeps the sums of those
// three bins in registers and adds a bin into the block's LDS copy of the
// projection (a double atomic) once the column has left it.  The terms go
// into the window with fused multiply-adds by 0 or 1, exact products, so
// each term is added by one rounding.  Beside each sum, W accumulates |pixel|
// over every pixel whose window holds the bin: each term is at most
// |pixel| / 4 (1 + 2^-39), so W bounds the bin's sum|terms|.
// Output per plane: sums at [a][r], bounds at [4 + a][r].
template <int NA>
__global__ void __launch_bounds__(256) k_radon_sum(const Img *imgs, const double *trig, double *const *proj, int ngrp,
                                                   int nimg) {
  extern __shared__ double lds[];
  int ent, grp;
  if (!xcd_item(ngrp, nimg, ent, grp)) return;
  const Img im = imgs[ent];
  const int M = im.h, N = im.w;
  const int rLast = (int)ceil(sqrt((double)((N - 1 - (N - 1) / 2) * (N - 1 - (N - 1) / 2)) +
                                   (double)((M - 1 - (M - 1) / 2) * (M - 1 - (M - 1) / 2)))) + 1;
  const int rFirst = -rLast, rSize = 2 * rLast + 1;
  const int a0 = grp * NA;
  double *S = lds, *W = lds + NA * rSize;
  for (int i = threadIdx.x; i < 2 * NA * rSize; i += blockDim.x) lds[i] = 0.0;
  __syncthreads();
  const int xO = (N - 1) / 2, yO = (M - 1) / 2, M2 = 2 * M;
  double cs[NA], sn[NA];
#pragma unroll
  for (int a = 0; a < NA; a++) {
    cs[a] = trig[2 * (a0 + a)];
    sn[a] = trig[2 * (a0 + a) + 1];
  }
  const gptr<const double> P = to_global((const double *)im.p);
  const double c0 = -yO - 0.25;
  // a bin's sum into LDS; w == 0 means no nonzero pixel reached it (s == +0)
  auto flush = [&](int a, int b, double sv, double wv) {
    if (wv != 0.0) {
      b = min(max(b, 0), rSize - 1);
      atomicAdd(&S[a * rSize + b], sv);
      atomicAdd(&W[a * rSize + b], wv);
    }
  };
  for (int c = threadIdx.x; c < N; c += blockDim.x) {
    double xa[NA], xb[NA], R0[NA], R1[NA], R2[NA], A0[NA], A1[NA], A2[NA];
    double base[NA];  // the window's first bin
#pragma unroll
    for (int a = 0; a < NA; a++) {
      xa[a] = radon_xc(2 * c, xO, cs[a], rFirst);
      xb[a] = radon_xc(2 * c + 1, xO, cs[a], rFirst);
      R0[a] = R1[a] = R2[a] = A0[a] = A1[a] = A2[a] = 0.0;
      base[a] = 1e9;  // above every bin: the first pixel opens the window
    }
    double pn = P[c];
    for (int y = 0; y < M; y++) {
      const double pix = pn;
      if (y + 1 < M) pn = P[(long)(y + 1) * N + c];
      const double q = pix * 0.25, ap = fabs(pix);
      // yTable of sub-rows 2y and 2y+1 (exact multiples of 0.25)
      const double ya = c0 + 0.5 * (M2 - 1 - 2 * y), yb = ya - 0.5;
#pragma unroll
      for (int a = 0; a < NA; a++) {
        const double ysa = ya * sn[a], ysb = yb * sn[a];
        const double r00 = xa[a] + ysa, r01 = xa[a] + ysb, r10 = xb[a] + ysa, r11 = xb[a] + ysb;
        // rLow = (int)rIdx as a double (rIdx >= 0 is far below 2^31)
        const double l00 = trunc(r00), l01 = trunc(r01), l10 = trunc(r10), l11 = trunc(r11);
        const double pb = fmin(fmin(l00, l01), fmin(l10, l11));
        if (pb < base[a]) {
          if (base[a] - pb >= 3.0) {
            flush(a, (int)base[a], R0[a], A0[a]);
            flush(a, (int)base[a] + 1, R1[a], A1[a]);
            flush(a, (int)base[a] + 2, R2[a], A2[a]);
            R0[a] = R1[a] = R2[a] = A0[a] = A1[a] = A2[a] = 0.0;
          } else {
            do {
              flush(a, (int)base[a] + 2, R2[a], A2[a]);
              R2[a] = R1[a]; R1[a] = R0[a]; R0[a] = 0.0;
              A2[a] = A1[a]; A1[a] = A0[a]; A0[a] = 0.0;
              base[a] -= 1.0;
            } while (pb < base[a]);
          }
          base[a] = pb;
        }
        auto add = [&](double rIdx, double rLow) {
          const double pixelLow = q * (1 - rIdx + rLow), pixelHigh = q - pixelLow;
          const double d = rLow - pb, e = 1.0 - d;  // rLow - pb is 0 or 1
          R0[a] = fma(e, pixelLow, R0[a]);
          R1[a] = fma(d, pixelLow, R1[a]);
          R1[a] = fma(e, pixelHigh, R1[a]);
          R2[a] = fma(d, pixelHigh, R2[a]);
        };
        add(r00, l00);
        add(r01, l01);
        add(r10, l10);
        add(r11, l11);
        A0[a] += ap;
        A1[a] += ap;
        A2[a] += ap;
      }
    }
#pragma unroll
    for (int a = 0; a < NA; a++) {
      flush(a, (int)base[a], R0[a], A0[a]);
      flush(a, (int)base[a] + 1, R1[a], A1[a]);
      flush(a, (int)base[a] + 2, R2[a], A2[a]);
    }
  }
  __syncthreads();
  double *out = proj[ent];
  for (int i = threadIdx.x; i < NA * rSize; i += blockDim.x) {
    out[(long)a0 * rSize + i] = S[i];
    out[(long)(4 + a0) * rSize + i] = W[i];
  }
}

// Per plane (block of 256): wave a takes angle a.  The bins whose interval
// [sum - d, sum + d], d = 4 gamma_K W slack, reaches the smallest upper end
// or the largest lower end are the only ones that can be the minimum or the
// maximum; they are recomputed in the reference's order, which gives the
// reference's hmin and hmax exactly.  A bin's slot is then decided from its
// interval (hist_bin is monotone in the value) or, when the interval straddles
// a slot boundary, from its recomputed value.  K = 4MN + 4M + N + 8 bounds the
// number of additions on any path of either summation tree.  Non-finite sums
// take the reference's path for that angle: every bin recomputed, the
// histogram formed serially as radon.cpp does.
__global__ void __launch_bounds__(256) k_radon_cert(const Img *imgs, const double *trig, double *const *proj,
                                                    OutRef o, double slack) {
  __shared__ RadonWaveLds wl[4];
  __shared__ double v[12];  // the 12 counts, angle-major
  const int a = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const Img im = imgs[blockIdx.x];
  const int M = im.h, N = im.w;
  const int rLast = (int)ceil(sqrt((double)((N - 1 - (N - 1) / 2) * (N - 1 - (N - 1) / 2)) +
                                   (double)((M - 1 - (M - 1) / 2) * (M - 1 - (M - 1) / 2)))) + 1;
  const int rFirst = -rLast, rSize = 2 * rLast + 1;
  const double K = 4.0 * M * N + 4.0 * M + N + 8.0, u = 0x1p-53;
  const double cb = 4.0 * (K * u / (1.0 - K * u)) * slack;
  // sums of subnormal-sized terms: a floor keeps d from underflowing to 0
  auto dlt = [&](double wv) { return wv > 0.0 ? fmax(cb * wv, 0x1p-1000) : 0.0; };
  const gptr<const double> P = to_global((const double *)im.p);
  double *const p = proj[blockIdx.x] + (long)a * rSize, *const w = proj[blockIdx.x] + (long)(4 + a) * rSize;
  const double cs = trig[2 * a], sn = trig[2 * a + 1];
  double L = __builtin_inf(), U = -__builtin_inf();
  bool nonfinite = false;
  for (int r = lane; r < rSize; r += 64) {
    const double d = dlt(w[r]), lo = p[r] - d, hi = p[r] + d;
    if (!(fabs(lo) <= 1.7976931348623157e308 && fabs(hi) <= 1.7976931348623157e308)) nonfinite = true;
    L = fmin(L, hi);
    U = fmax(U, lo);
  }
  L = wave_all_reduce(L, MinOp());
  U = wave_all_reduce(U, MaxOp());
  const bool bad = __ballot(nonfinite) != 0;
  // candidates (every bin when bad), each recomputed by the whole wave
  for (int c0 = 0; c0 < rSize; c0 += 64) {
    const int r = c0 + lane;
    bool need = false;
    if (r < rSize) {
      const double wv = w[r];
      if (bad ? wv != 0.0 : wv > 0.0) {
        const double d = dlt(wv), lo = p[r] - d, hi = p[r] + d;
        need = bad || lo <= L || hi >= U;
      }
    }
    for (unsigned long long m = __ballot(need); m; m &= m - 1) {
      const int k = __ffsll((long long)m) - 1;
      const double e = radon_bin_wave(P, M, N, cs, sn, rFirst, c0 + k, wl[a]);
      if (lane == k) {
        p[r] = e;
        w[r] = 0.0;
      }
    }
  }
  wave_sync();
  if (bad) {  // every bin is the reference's: its serial histogram
    if (lane == 0) {
      double hmin = kInf, hmax = -kInf;
      for (int i = 0; i < rSize; i++) { if (p[i] > hmax) hmax = p[i]; if (p[i] < hmin) hmin = p[i]; }
      const double hs = hist_scale(hmin, hmax, 3);
      double b0 = 0, b1 = 0, b2 = 0;
      for (int i = 0; i < rSize; i++) {
        const int b = hist_bin(p[i], hmin, hs, 3);
        b0 += b == 0 ? 1.0 : 0.0;
        b1 += b == 1 ? 1.0 : 0.0;
        b2 += b == 2 ? 1.0 : 0.0;
      }
      v[3 * a] = b0;
      v[3 * a + 1] = b1;
      v[3 * a + 2] = b2;
    }
  } else {
    // the minimum and maximum bins are among the recomputed (or exact) ones
    double mn = __builtin_inf(), mx = -__builtin_inf();
    for (int r = lane; r < rSize; r += 64) {
      if (w[r] == 0.0) {
        mn = fmin(mn, p[r]);
        mx = fmax(mx, p[r]);
      }
    }
    mn = wave_all_reduce(mn, MinOp());
    mx = wave_all_reduce(mx, MaxOp());
    const double hs = hist_scale(mn, mx, 3);
    int n0 = 0, n1 = 0, n2 = 0;
    for (int c0 = 0; c0 < rSize; c0 += 64) {
      const int r = c0 + lane;
      int b = -1;
      bool open = false;
      if (r < rSize) {
        const double wv = w[r], pv = p[r];
        if (wv == 0.0) {
          b = hist_bin(pv, mn, hs, 3);
        } else {
          const double d = dlt(wv);
          b = hist_bin(fmax(pv - d, mn), mn, hs, 3);
          open = hist_bin(fmin(pv + d, mx), mn, hs, 3) != b;
        }
      }
      for (unsigned long long m = __ballot(open); m; m &= m - 1) {
        const int k = __ffsll((long long)m) - 1;
        const double e = radon_bin_wave(P, M, N, cs, sn, rFirst, c0 + k, wl[a]);
        if (lane == k) b = hist_bin(e, mn, hs, 3);
      }
      n0 += b == 0;
      n1 += b == 1;
      n2 += b == 2;
    }
    n0 = wave_all_reduce(n0, SumOp());
    n1 = wave_all_reduce(n1, SumOp());
    n2 = wave_all_reduce(n2, SumOp());
    if (lane == 0) {
      v[3 * a] = n0;
      v[3 * a + 1] = n1;
      v[3 * a + 2] = n2;
    }
  }
  __syncthreads();
  // output order of vd_RadonTextures (radon.cpp:98-115): angles 0, 135, 45, 90
  if (threadIdx.x < 12) {
    const int t = threadIdx.x, src = t < 3 ? t : t < 6 ? t + 6 : t - 3;
    o.rows[blockIdx.x][t] = v[src];
  }
}

// rSize above which the certified form's LDS (16 bytes per bin and angle)
// does not fit one angle per block
constexpr int RADON_CERT_MAXBINS = 10240;

bool radon_certified_fits(int maxbins) { return maxbins <= RADON_CERT_MAXBINS; }

void launch_radon_certified(const Img *imgs, const double *trig, double *const *proj, OutRef o, int nimg, int maxbins,
                            double slack, hipStream_t s) {
  // angles per block: as many as keep the block's LDS near 48 KB
  const int na = (64 * maxbins <= 49152) ? 4 : (32 * maxbins <= 49152) ? 2 : 1;
  const size_t lds = (size_t)2 * na * maxbins * sizeof(double);
  static const bool lds_ok =
      hipFuncSetAttribute((const void *)k_radon_sum<1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          2 * RADON_CERT_MAXBINS * (int)sizeof(double)) == hipSuccess;
  (void)lds_ok;
  const int ngrp = 4 / na;
  if (na == 4)
    k_radon_sum<4><<<xcd_grid(ngrp, nimg), 256, lds, s>>>(imgs, trig, proj, ngrp, nimg);
  else if (na == 2)
    k_radon_sum<2><<<xcd_grid(ngrp, nimg), 256, lds, s>>>(imgs, trig, proj, ngrp, nimg);
  else
    k_radon_sum<1><<<xcd_grid(ngrp, nimg), 256, lds, s>>>(imgs, trig, proj, ngrp, nimg);
  k_radon_cert<<<nimg, 256, 0, s>>>(imgs, trig, proj, o, slack);
}

void launch_radon(const Img *imgs, const double *trig, double *const *proj, OutRef o, int nimg, int maxbins,
                  int maxh, int maxw, hipStream_t s) {
  (void)maxw;
  // LDS: the y*sin table (2 maxh <= 8192 doubles), then, for the slanted
  // angles, the chunk buffer
  static const bool lds_ok = hipFuncSetAttribute((const void *)k_radon_proj, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                  (2 * 4096 + RADON_TILE) * (int)sizeof(double)) == hipSuccess;
  (void)lds_ok;
  const int ysz = 2 * maxh;
  const int nbg = (maxbins + 255) / 256;
  // angle 0 (direct reads, small LDS for occupancy), then angles 1..3
  k_radon_proj<<<xcd_grid(nbg, nimg), 256, (size_t)ysz * sizeof(double), s>>>(imgs, trig, proj, nbg, nimg, ysz, 0, 1);
  k_radon_proj<<<xcd_grid(3 * nbg, nimg), 256, ((size_t)ysz + RADON_TILE) * sizeof(double), s>>>(imgs, trig, proj, nbg,
                                                                                                 nimg, ysz, 1, 3);
  k_radon_hist<<<nimg, 64, 0, s>>>(imgs, proj, o);
}

}  // namespace wnd
