// Otsu / Inverse-Otsu object features, Prewitt edge features and Gabor
// textures.  FP64.  Integer outputs (thresholds, labels, areas, counts,
// Euler number) are exact; the connected components are canonicalised to
// the reference's raster-order labelling.
#include "common.hpp"
#include "kernels.hpp"
#include "blocksort.hpp"

namespace wnd {

// Otsu threshold of a plane (ImageMatrix::Otsu(dynamic), cmatrix.cpp:1520-1570):
// 1024-bin histogram over [mn, mx], then the reference's omega / myu / sigma
// recurrence.  The threshold of a filled 1024-bin histogram, by the whole
// block (every thread calls it; uniform result): the bin fractions
// hist[i] / wh and, once thread 0 has run the sequential prefix sums om / my
// (and myu[L-1]) with the reference's operations, each bin's between-class
// variance are formed in parallel; the threshold is the first bin of the
// largest variance above 0 (the reference's strict comparison from 0).
__device__ double otsu_threshold(const int *hist, long n, double mn, double mx) {
  __shared__ double pf[1024], om_s[1024], my_s[1024];
  __shared__ double red_d[32];
  __shared__ long long red_l[32];
  __shared__ double my_last_s;
  const double wh = (double)(unsigned)n;
  for (int i = threadIdx.x; i < 1024; i += blockDim.x) pf[i] = hist[i] / wh;
  __syncthreads();
  if (threadIdx.x == 0) {
    // myu[L-1] first (the recurrence below repeats the same operations)
    double my_last = 0.0;
    for (int i = 1; i < 1024; i++) my_last = my_last + i * pf[i];
    my_last_s = my_last;
    double om = pf[0], my = 0.0;
    for (int i = 0; i < 1023; i++) {
      if (i > 0) {
        om = om + pf[i];
        my = my + i * pf[i];
      }
      om_s[i] = om;
      my_s[i] = my;
    }
  }
  __syncthreads();
  const double my_last = my_last_s;
  double best = 0.0;
  long long bi = 1023;  // no variance above 0: threshold bin 0 (below)
  for (int i = threadIdx.x; i < 1023; i += blockDim.x) {
    const double om = om_s[i], my = my_s[i];
    double sg;
    if (om != 0.0 && om != 1.0) {
      const double d = my_last * om - my;
      sg = (d * d) / (om * (1.0 - om));
    } else sg = 0.0;
    if (sg > best) { best = sg; bi = i; }  // this thread's bins ascend: its first maximum
  }
  const double top = block_reduce(best, MaxOp(), red_d);
  long long cand = (top > 0.0 && best == top) ? bi : 1023;
  cand = block_reduce(cand, MinOp(), red_l);
  const int thr = top > 0.0 ? (int)cand : 0;
  return ((((double)thr / (double)(1023)) * (mx - mn)) + mn);
}

template <typename ValF>
__device__ double block_otsu(long n, double mn, double mx, ValF val, int *hist) {
  for (int b = threadIdx.x; b < 1024; b += blockDim.x) hist[b] = 0;
  __syncthreads();
  const double sc = hist_scale(mn, mx, 1024);
#pragma unroll 4
  for (long i = threadIdx.x; i < n; i += blockDim.x) atomicAdd(&hist[hist_bin(val(i), mn, sc, 1024)], 1);
  __syncthreads();
  return otsu_threshold(hist, n, mn, mx);
}

// ------------------------------------------------------------- union-find
__device__ __forceinline__ int uf_find(volatile int *lab, int x) {
  while (true) {
    const int p = lab[x];
    if (p == x) return x;
    x = p;
  }
}
__device__ void uf_union(int *lab, int a, int b) {
  while (true) {
    a = uf_find(lab, a);
    b = uf_find(lab, b);
    if (a == b) return;
    if (a < b) { const int t = a; a = b; b = t; }
    const int old = atomicMin(&lab[a], b);
    if (old == a) return;
    a = old;
  }
}

// Object features (ImageMatrix::FeatureStatistics, cmatrix.cpp:1642-1717;
// bwlabel / GlobalCentroid / FeatureCentroid / EulerNumber,
// FeatureStatistics.cpp:44-254; output layout FeatureAlgorithms.cpp:370-394).
// One 1024-thread work-group per image.  invert != 0 evaluates the
// Inverse-Otsu variant on v' = (max - v) + min (ImageMatrix::invert, cmatrix.cpp:569-572).
__global__ void __launch_bounds__(1024) k_objects(const Img *imgs, const Stats *st, int invert, const ObjDesc *ods,
                                                  OutRef o) {
  __shared__ int hist[1024];
  __shared__ double shd[16];
  __shared__ long long shl[16];
  __shared__ int scan_w[16];
  __shared__ int s_base;
  const Img im = imgs[blockIdx.x];
  const ObjDesc OD = ods[blockIdx.x];
  const Stats S = st[blockIdx.x];
  const int W = im.w, H = im.h;
  const long n = (long)W * H;
  const int t = threadIdx.x;
  const double smx = S.mx, smn = S.mn;
  auto val = [&](long i) -> double { return invert ? (smx - im.p[i]) + smn : im.p[i]; };
  // stats of the plane the features are taken on
  double mn = smn, mx = smx;
  if (invert) {
    double a = 1e308, b = -1e308;
    for (long i = t; i < n; i += 1024) {
      const double v = val(i);
      if (moments_ok(v)) { a = fmin(a, v); b = fmax(b, v); }
    }
    mn = block_reduce(a, MinOp(), shd);
    mx = block_reduce(b, MaxOp(), shd);
  }
  const double thr = block_otsu(n, mn, mx, val, hist);
  int *lab = OD.label;
  // mask, labels, global centroid (exact integer sums).  A wave takes 64
  // consecutive pixels (raster order); each mask pixel starts at the first
  // pixel of its run inside that segment (a ballot), so runs need no unions
  // within a segment and the trees stay shallow.
  long long xm = 0, ym = 0, mass = 0, c1 = 0, c3 = 0;
  const int lane = t & 63;
  for (long c0 = 0; c0 < n; c0 += 1024) {
    const long i = c0 + t;
    const bool m = i < n && val(i) > thr;
    const int y = (int)(i / W), x = (int)(i - (long)y * W);
    // bit l: pixel l of the segment continues the run of pixel l - 1 (both
    // in the mask, the same row)
    const unsigned long long mb = __ballot(m);
    const unsigned long long cont = __ballot(m && x > 0) & (mb << 1);
    const unsigned long long below = lane == 63 ? ~0ull : ((2ull << lane) - 1);
    const int head = 63 - __clzll(~cont & below);
    if (i < n) lab[i] = m ? (int)(i - (lane - head)) : -1;
    if (m) {
      xm += x + 1; ym += y + 1; mass++;
    }
  }
  __syncthreads();
  // 8-connected union-find; every root ends at the component's smallest
  // index.  Of a horizontal run [s, e] the upper neighbours are s-1 .. e+1:
  // each pixel joins its upper-right one, the run's first pixel also its
  // upper-left and upper ones; a segment's first pixel joins its left
  // neighbour when the run crosses the segment boundary.
  for (long i = t; i < n; i += 1024) {
    if (lab[i] < 0) continue;
    const int y = i / W, x = i - (long)y * W;
    const bool left = x > 0 && lab[i - 1] >= 0;
    if (left && (i & 63) == 0) uf_union(lab, (int)i, (int)(i - 1));
    if (y > 0) {
      const long u = i - W;
      if (x < W - 1 && lab[u + 1] >= 0) uf_union(lab, (int)i, (int)(u + 1));
      if (!left) {
        if (x > 0 && lab[u - 1] >= 0) uf_union(lab, (int)i, (int)(u - 1));
        if (lab[u] >= 0) uf_union(lab, (int)i, (int)u);
      }
    }
  }
  __syncthreads();
  __threadfence_block();
  for (long i = t; i < n; i += 1024)
    if (lab[i] >= 0) lab[i] = uf_find(lab, (int)i);
  __syncthreads();
  // Euler number patterns (EulerNumber(8): (C1 - C3) / 4; Cd is never counted)
  for (long i = t; i < n; i += 1024) {
    const int y = i / W, x = i - (long)y * W;
    if (y < 1 || x < 1) continue;
    const int c = (lab[i - W - 1] >= 0) + (lab[i - W] >= 0) + (lab[i - 1] >= 0) + (lab[i] >= 0);
    c1 += (c == 1);
    c3 += (c == 3);
  }
  xm = block_reduce(xm, SumOp(), shl);
  ym = block_reduce(ym, SumOp(), shl);
  mass = block_reduce(mass, SumOp(), shl);
  c1 = block_reduce(c1, SumOp(), shl);
  c3 = block_reduce(c3, SumOp(), shl);
  // label ids: rank of each root in raster order (1-based)
  int *rid = (int *)OD.keys;  // reuse: rid[i] for roots, before the keys are needed
  if (t == 0) s_base = 0;
  __syncthreads();
  for (long c0 = 0; c0 < n; c0 += 1024) {
    const long i = c0 + t;
    const int isroot = (i < n && lab[i] == (int)i) ? 1 : 0;
    // block exclusive scan of isroot
    const int wv = t >> 6;
    const uint64_t bal = __ballot(isroot);
    const int pre = __popcll(bal & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
    if (lane == 0) scan_w[wv] = __popcll(bal);
    __syncthreads();
    int woff = 0, tot = 0;
    for (int w = 0; w < 16; w++) { if (w < wv) woff += scan_w[w]; tot += scan_w[w]; }
    if (isroot) rid[i] = s_base + woff + pre + 1;
    __syncthreads();
    if (t == 0) s_base += tot;
    __syncthreads();
  }
  const int count = s_base;
  int *area = OD.area;
  long long *axm = OD.xm, *aym = OD.ym;
  for (long l = t; l <= count; l += 1024) { area[l] = 0; axm[l] = 0; aym[l] = 0; }
  __syncthreads();
  // areas and coordinate sums per object: the lanes of a wave holding one
  // label add theirs first, then one lane adds the totals (exact integers)
  for (long c0 = 0; c0 < n; c0 += 1024) {
    const long i = c0 + t;
    int l = -1;
    long long px = 0, py = 0;
    if (i < n) {
      const int r = lab[i];
      if (r >= 0) {
        l = rid[r];
        const int y = i / W, x = i - (long)y * W;
        px = x + 1;
        py = y + 1;
      }
    }
    for (unsigned long long pend = __ballot(l >= 0); pend;) {
      const int lead = __ffsll((long long)pend) - 1;
      const int ll = __shfl(l, lead, 64);
      const bool mine = l == ll;
      const unsigned long long mm = __ballot(mine);
      long long sx = mine ? px : 0, sy = mine ? py : 0;
      for (int off = 32; off > 0; off >>= 1) {
        sx += __shfl_xor(sx, off, 64);
        sy += __shfl_xor(sy, off, 64);
      }
      if (lane == lead) {
        atomicAdd(&area[ll], (int)__popcll(mm));
        atomicAdd((unsigned long long *)&axm[ll], (unsigned long long)sx);
        atomicAdd((unsigned long long *)&aym[ll], (unsigned long long)sy);
      }
      pend &= ~mm;
    }
  }
  __syncthreads();
  // per-object distance to the global centroid
  double cx = 0, cy = 0;
  if (mass) { cx = (double)xm / (double)mass; cy = (double)ym / (double)mass; }
  uint64_t *ka = OD.keys, *kb = OD.keys + (count + 1);
  double *dist = OD.dist;
  double sum_d = 0;
  for (long l = t; l < count; l += 1024) {
    const long L = l + 1;
    const double xc = (double)axm[L] / (double)area[L], yc = (double)aym[L] / (double)area[L];
    const double dx = xc - cx, dy = yc - cy;
    const double d = sqrt(dx * dx + dy * dy);
    dist[l] = d;
    sum_d += d;
  }
  sum_d = block_reduce(sum_d, SumOp(), shd);
  // sort the areas (as integer keys)
  for (long l = t; l < count; l += 1024) ka[l] = (uint64_t)area[l + 1];
  __syncthreads();
  uint64_t *sa = block_radix_sort(ka, kb, count);
  // sorted areas -> dist buffer tail? keep them in place: copy to area[] (1-based free now)
  for (long l = t; l < count; l += 1024) area[l] = (int)sa[l];
  __syncthreads();
  for (long l = t; l < count; l += 1024) ka[l] = dkey(dist[l]);
  __syncthreads();
  uint64_t *sd = block_radix_sort(ka, kb, count);
  for (long l = t; l < count; l += 1024) dist[l] = dkey_inv(sd[l]);
  __syncthreads();
  // statistics (count == 0 reads an empty array in the reference; defined as 0 here)
  const double sum_a = (double)mass;  // sum of object areas == mask pixel count
  double AreaMean = 0, DistMean = 0, amin = 0, amax = 0, dmin = 0, dmax = 0;
  if (count > 0) {
    amin = area[0]; amax = area[count - 1];
    dmin = dist[0]; dmax = dist[count - 1];
    AreaMean = sum_a / count;
    DistMean = sum_d / count;
  }
  if (t < 20) hist[t] = 0;
  __syncthreads();
  double va = 0, vd = 0;
  const double hsa = (amax - amin > 0) ? 10.0 / (double)(amax - amin) : 0.0;
  const double hsd = (dmax - dmin > 0) ? 10.0 / (double)(dmax - dmin) : 0.0;
  for (long l = t; l < count; l += 1024) {
    if (amax - amin > 0) {
      const double da = area[l] - AreaMean;
      va += da * da;
      atomicAdd(&hist[hist_bin((double)area[l], amin, hsa, 10)], 1);
    }
    const double dd = dist[l] - DistMean;
    vd += dd * dd;
    atomicAdd(&hist[10 + hist_bin(dist[l], dmin, hsd, 10)], 1);
  }
  va = block_reduce(va, SumOp(), shd);
  vd = block_reduce(vd, SumOp(), shd);
  if (t == 0) {
    double *out = o.rows[blockIdx.x];
    int k = 0;
    for (int j = 0; j < 10; j++) out[k++] = hist[j];
    out[k++] = amax;
    out[k++] = AreaMean;
    out[k++] = count > 0 ? (double)(unsigned)area[count / 2] : 0.0;
    out[k++] = amin;
    out[k++] = count > 1 ? va / (count - 1) : va;
    out[k++] = cx;
    out[k++] = cy;
    out[k++] = count;
    for (int j = 0; j < 10; j++) out[k++] = hist[10 + j];
    out[k++] = dmax;
    out[k++] = DistMean;
    out[k++] = count > 0 ? dist[count / 2] : 0.0;
    out[k++] = dmin;
    out[k++] = count > 1 ? vd / (count - 1) : vd;
    out[k++] = (double)((c1 - c3) / 4);
  }
}

void launch_objects(const Img *imgs, const Stats *st, int invert, const ObjDesc *od, OutRef o, int nimg, int maxn,
                    hipStream_t s) {
  (void)maxn;
  k_objects<<<nimg, 1024, 0, s>>>(imgs, st, invert, od, o);
}

// ================================================================ Edge features
// Prewitt magnitude / direction (cmatrix.cpp:1317-1384), x outer order of the
// reference is irrelevant per pixel; the accumulation order inside a pixel is kept.
__global__ void k_prewitt(const Img *in, const Img *mag, const Img *dir) {
  const Img a = in[blockIdx.y], om = mag[blockIdx.y], od = dir[blockIdx.y];
  const long w = a.w, h = a.h;
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= w * h) return;
  const long y = i / w, x = i - y * w;
  double sumx = 0, sumy = 0;
  for (long j = y - 1; j <= y + 1; j++) if (j >= 0 && j < h && x - 1 >= 0) sumx += a.p[j * w + x - 1] * 1;
  for (long j = y - 1; j <= y + 1; j++) if (j >= 0 && j < h && x + 1 < w) sumx += a.p[j * w + x + 1] * -1;
  for (long k = x - 1; k <= x + 1; k++) if (k >= 0 && k < w && y - 1 >= 0) sumy += a.p[(y - 1) * w + k] * 1;
  for (long k = x - 1; k <= x + 1; k++) if (k >= 0 && k < w && y + 1 < h) sumy += a.p[(y + 1) * w + k] * -1;
  om.p[i] = sqrt(sumx * sumx + sumy * sumy);
  double d;
  if (sumy == 0 || fabs(sumy) < 1 / kInf) d = 3.1415926 * (sumx < 0 ? 1 : 0);
  else d = atan2(sumy, sumx);
  od.p[i] = d;
}

// EdgeStatistics (cmatrix.cpp:1402-1458) + output layout (FeatureAlgorithms.cpp:314-335).
// st2[2*i] / st2[2*i+1] and medians[2*i] / [2*i+1] belong to magnitude / direction.
__global__ void __launch_bounds__(256) k_edge_final(const Img *mag, const Img *dir, const Stats *st2,
                                                    const double *medians, OutRef o) {
  __shared__ long long shl[4];
  __shared__ int hm[8], hd[8];
  const Img M = mag[blockIdx.x], D = dir[blockIdx.x];
  const Stats SM = st2[2 * blockIdx.x], SD = st2[2 * blockIdx.x + 1];
  const long n = (long)M.h * M.w;
  if (threadIdx.x < 8) { hm[threadIdx.x] = 0; hd[threadIdx.x] = 0; }
  __syncthreads();
  const double level = SM.mean;
  const double sm = hist_scale(SM.mn, SM.mx, 8), sd = hist_scale(SD.mn, SD.mx, 8);
  long long area = 0;
  for (long i = threadIdx.x; i < n; i += blockDim.x) {
    const double m = M.p[i];
    if (m > level) area++;
    atomicAdd(&hm[hist_bin(m, SM.mn, sm, 8)], 1);
    atomicAdd(&hd[hist_bin(D.p[i], SD.mn, sd, 8)], 1);
  }
  area = block_reduce(area, SumOp(), shl);
  if (threadIdx.x == 0) {
    double DH[8], MH[8], diff[4] = {0, 0, 0, 0}, homog = 0;
    for (int b = 0; b < 8; b++) { DH[b] = hd[b]; MH[b] = hm[b]; }
    for (int b = 0; b < 4; b++) diff[b] = fabs(DH[b] - DH[b + 4]);
    double sum = 0;
    for (int b = 0; b < 4; b++) {
      if (DH[b] + DH[b + 4] != 0) diff[b] = diff[b] / (DH[b] + DH[b + 4]);
      sum += (DH[b] + DH[b + 4]);
    }
    if (sum > 0) homog = (DH[0] + DH[1]) / sum;
    double *c = o.rows[blockIdx.x];
    int h = 0;
    c[h++] = (double)area;
    for (int j = 0; j < 4; j++) c[h++] = diff[j];
    for (int j = 0; j < 8; j++) c[h++] = DH[j];
    c[h++] = homog;
    c[h++] = SD.mean;
    c[h++] = medians[2 * blockIdx.x + 1];
    c[h++] = SD.var;
    for (int j = 0; j < 8; j++) c[h++] = MH[j];
    c[h++] = SM.mean;
    c[h++] = medians[2 * blockIdx.x];
    c[h++] = SM.var;
  }
}

void launch_prewitt(const Img *imgs, const Img *mag, const Img *dir, int nimg, long maxn, hipStream_t s) {
  k_prewitt<<<dim3((maxn + 255) / 256, nimg), 256, 0, s>>>(imgs, mag, dir);
}

void launch_edge_final(const Img *mag, const Img *dir, const Stats *st2, const double *medians, OutRef o, int nimg,
                       hipStream_t s) {
  k_edge_final<<<nimg, 256, 0, s>>>(mag, dir, st2, medians, o);
}

// ======================================================================= Gabor
// GaborTextureFilters2D (gabor.cpp:201-231) with GaborEnergy / conv2comp
// (gabor.cpp:19-52,166-195).  kernels: 8 host-computed normalised complex
// 38x38 filters (index 0 = low-pass f0 = 0.1, 1..7 = f0 1..7).  Each output
// pixel accumulates the 1444 taps in the reference's (row, column) tap order
// with separate multiply and add; taps that fall outside the image add
// +-0, which leaves the sum unchanged.
#define GB_N 38
#define GB_TR 32          // output rows per block
#define GB_TC 64          // output columns per block
#define GB_LD (GB_TC + GB_N - 1 + 1)
// Block = 32 x 64 outputs of one filter on one image; thread = 8 adjacent
// outputs of one row.  Taps are taken in groups of eight: the 15 input values
// a group needs for the thread's 8 outputs are read from LDS once into
// registers, and every output still adds its taps in the reference's (j, i)
// order (conv2comp, gabor.cpp:19-52).
//   CERT = false: separate multiply and add, the reference's rounding.
//   CERT = true: fused multiply-adds (half the instructions), and beside
// each energy E an a-priori bound D on its distance from the CERT = false
// value E64, for k_gabor_post_exact to certify every count (k_gabor_repair
// recomputes in the CERT = false order each energy a count leaves open).
// Both sums of a component are within gamma_1444 S of the exact sum,
// S = sum |x| |w| <= (largest |x| of the block's input tile) x wsum_f, and
// the two magnitudes round by at most 2u each, so |E - E64| <= 2 gamma_1444
// S + 4u E: D = coef * tmax * wsum_f + 1e-14 E, coef = 4e-13 (2 gamma_1444
// = 3.21e-13) times the context's slack, rounded up.
template <bool CERT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) k_gabor_conv_t(const Img *imgs, const double *kernels, double *const *energy,
                                                      const double *wsum, float *const *bound, double coef) {
  __shared__ double tile[GB_TR + GB_N - 1][GB_LD];
  __shared__ double2 wk[GB_N * GB_N];
  __shared__ double shd[4];
  const Img im = imgs[blockIdx.z];
  const int f = blockIdx.y % 8;
  const int tilesx = (im.w + GB_TC - 1) / GB_TC, tilesy = (im.h + GB_TR - 1) / GB_TR;
  if ((int)blockIdx.x >= tilesx * tilesy) return;
  const int b0 = (blockIdx.x / tilesx) * GB_TR, a0 = (blockIdx.x % tilesx) * GB_TC;
  const int t = threadIdx.x;
  const double *K = kernels + (long)f * GB_N * GB_N * 2;
  for (int e = t; e < GB_N * GB_N; e += 256) wk[e] = make_double2(K[2 * e], K[2 * e + 1]);
  // input rows b0-18 .. b0+GB_TR+18, columns a0-18 .. a0+GB_TC+18
  double tmax = 0.0;
  for (int r = t / 64; r < GB_TR + GB_N - 1; r += 4) {
    const int y = b0 - 18 + r;
    for (int c = t % 64; c < GB_TC + GB_N - 1; c += 64) {
      const int x = a0 - 18 + c;
      const double v = (y >= 0 && y < im.h && x >= 0 && x < im.w) ? im.p[(long)y * im.w + x] : 0.0;
      tile[r][c] = v;
      if (CERT) tmax = fmax(tmax, fabs(v));
    }
  }
  if (CERT) tmax = block_reduce(tmax, MaxOp(), shd);  // (synchronises the block)
  else __syncthreads();
  const int ty = t / 8, tx = t % 8;
  const int b = b0 + ty, a = a0 + tx * 8;
  double re[8], imv[8];
#pragma unroll
  for (int q = 0; q < 8; q++) re[q] = imv[q] = 0.0;
  const int rb = ty + 37, cb = tx * 8 + 37;
#pragma unroll 1
  for (int j = 0; j < GB_N; j++) {
    const double *row = &tile[rb - j][0];
    const double2 *wr = &wk[j * GB_N];
#pragma unroll
    for (int i0 = 0; i0 < GB_N; i0 += 8) {
      // u[k] = row[cb - i0 - 7 + k]: value of output q, tap i is u[q - (i - i0) + 7]
      // taps in this group: D = min(8, GB_N - i0); slots k < 8 - D are unused
      const int D = (GB_N - i0 < 8) ? GB_N - i0 : 8;
      double u[15];
#pragma unroll
      for (int k = 0; k < 15; k++) u[k] = (k >= 8 - D) ? row[cb - i0 - 7 + k] : 0.0;
#pragma unroll
      for (int di = 0; di < 8; di++) {
        if (i0 + di < GB_N) {
          const double2 w = wr[i0 + di];
#pragma unroll
          for (int q = 0; q < 8; q++) {
            const double v = u[q - di + 7];
            if (CERT) {
              re[q] = fma(v, w.x, re[q]);
              imv[q] = fma(v, w.y, imv[q]);
            } else {
              re[q] += v * w.x;
              imv[q] += v * w.y;
            }
          }
        }
      }
    }
  }
  double *E = energy[blockIdx.z] + (long)f * im.h * im.w;
  if (b < im.h)
#pragma unroll
    for (int q = 0; q < 8; q++)
      if (a + q < im.w) {
        const double e = sqrt(re[q] * re[q] + imv[q] * imv[q]);
        E[(long)b * im.w + a + q] = e;
        if (CERT)
          bound[blockIdx.z][(long)f * im.h * im.w + (long)b * im.w + a + q] =
              (float)((coef * tmax * wsum[f] + 1e-14 * e) * 1.000001);
      }
}

__global__ void __launch_bounds__(1024) k_gabor_post(const Img *imgs, double *const *energy, OutRef o) {
  __shared__ int hist[1024];
  __shared__ double shd[16];
  __shared__ long long shl[16];
  const Img im = imgs[blockIdx.x];
  const long n = (long)im.h * im.w;
  const double *E0 = energy[blockIdx.x];
  double *out = o.rows[blockIdx.x];
  long long orig = 0;
  for (int f = 0; f < 8; f++) {
    const double *E = E0 + (long)f * n;
    // (the per-pixel loops are unrolled so that four loads are in flight)
    double mx = -1e308;
#pragma unroll 4
    for (long i = threadIdx.x; i < n; i += 1024) mx = fmax(mx, E[i]);
    mx = block_reduce(mx, MaxOp(), shd);
    if (f == 0) {
      long long c = 0;
      const double lvl = mx * 0.4;
#pragma unroll 4
      for (long i = threadIdx.x; i < n; i += 1024) c += E[i] > lvl;
      orig = block_reduce(c, SumOp(), shl);
      continue;
    }
    // e / max, its min / max, Otsu threshold, count above
    double a = 1e308, b = -1e308;
#pragma unroll 4
    for (long i = threadIdx.x; i < n; i += 1024) {
      const double v = E[i] / mx;
      if (moments_ok(v)) { a = fmin(a, v); b = fmax(b, v); }
    }
    a = block_reduce(a, MinOp(), shd);
    b = block_reduce(b, MaxOp(), shd);
    const double thr = block_otsu(n, a, b, [&](long i) { return E[i] / mx; }, hist);
    long long c = 0;
#pragma unroll 4
    for (long i = threadIdx.x; i < n; i += 1024) c += (E[i] / mx) > thr;
    c = block_reduce(c, SumOp(), shl);
    if (threadIdx.x == 0) out[f - 1] = (double)c / (double)orig;
  }
}

// The Gabor ratios from energies E with bounds D (|E - E64| <= D, E64 =
// k_gabor_conv_t<false>'s value: the fused multiply-add, FP32 and BF16
// forms), equal to k_gabor_post on the FP64 energies.  Every decision whose
// interval [E - D, E + D] leaves it open is taken on E64: the pixels
// concerned are marked (D < 0), and the repairs recompute them in
// k_gabor_conv_t<false>'s tap order (E = E64, D = 0).  Five phases of
// k_gabor_post_exact, one block per (image, filter), each followed by the
// repairs except the last:
//   1  mark the candidates for the maximum (upper end >= largest lower end)
//   2  max M; filter 0: mark the pixels open at 0.4 M; other filters: mark
//      the candidates for the minimum of e / M
//   3  filter 0: the base count; others: min a and max b of e / M, mark the
//      pixels whose Otsu bin is open
//   4  Otsu histogram and threshold; mark the pixels open at the threshold
//   5  the counts above the threshold, the ratios
// Per image and filter, params holds M, a, b, the threshold (and the base
// count in filter 0's slot).
constexpr int GP_STRIDE = 4;

// the marked pixels (D < 0) of one 32 x 64 tile of one filter's map: the
// input window staged in LDS as k_gabor_conv_t does, each marked output then
// summed in k_gabor_conv_t<false>'s tap order by one lane
__device__ void gabor_repair_tile(const Img &im, int tile, const double *K, double *E, float *D,
                                  double (*tile_px)[GB_LD], double2 *wk, int *list, int *nlist) {
  const int tilesx = (im.w + GB_TC - 1) / GB_TC;
  const int b0 = (tile / tilesx) * GB_TR, a0 = (tile % tilesx) * GB_TC;
  const int t = threadIdx.x;
  if (t == 0) *nlist = 0;
  __syncthreads();
  for (int e = t; e < GB_TR * GB_TC; e += 256) {
    const int r = e / GB_TC, c = e - r * GB_TC;
    const int y = b0 + r, x = a0 + c;
    if (y < im.h && x < im.w && D[(long)y * im.w + x] < 0.0f) list[atomicAdd(nlist, 1)] = e;
  }
  __syncthreads();
  const int nl = *nlist;
  if (nl == 0) return;  // (uniform: every thread returns)
  for (int e = t; e < GB_N * GB_N; e += 256) wk[e] = make_double2(K[2 * e], K[2 * e + 1]);
  for (int r = t / 64; r < GB_TR + GB_N - 1; r += 4) {
    const int y = b0 - 18 + r;
    for (int c = t % 64; c < GB_TC + GB_N - 1; c += 64) {
      const int x = a0 - 18 + c;
      tile_px[r][c] = (y >= 0 && y < im.h && x >= 0 && x < im.w) ? im.p[(long)y * im.w + x] : 0.0;
    }
  }
  __syncthreads();
  for (int k = t; k < nl; k += 256) {
    const int e = list[k];
    const int r = e / GB_TC, c = e - r * GB_TC;
    // output (b0 + r, a0 + c) reads tile row r + 37 - j, column c + 37 - i
    double re = 0.0, imv = 0.0;
    for (int j = 0; j < GB_N; j++) {
      const double *row = &tile_px[r + 37 - j][c + 37];
      const double2 *wr = &wk[j * GB_N];
#pragma unroll
      for (int i = 0; i < GB_N; i++) {
        const double v = row[-i];
        const double2 w = wr[i];
        re += v * w.x;
        imv += v * w.y;
      }
    }
    const long o = (long)(b0 + r) * im.w + a0 + c;
    E[o] = sqrt(re * re + imv * imv);
    D[o] = 0.0f;
  }
  __syncthreads();  // the buffers are reused by the next tile
}

// Per (image, filter), flags[(8 image + filter) * GR_STRIDE] holds the
// number of energies the last phase marked there and, when at most GR_CAP,
// their pixel indices.  k_gabor_repair_list recomputes listed energies one
// per thread straight from global memory; k_gabor_repair takes the filters
// with more, tile by tile.  Grid (tiles, 8, images), one tile per block, or
// (1, 8, images) with each block taking its filter's tiles in turn
// (loop_tiles: for bounds that leave few energies open).
constexpr int GR_CAP = 63;
constexpr int GR_STRIDE = GR_CAP + 1;
// One wavefront per (image, filter); per listed energy the lanes form the
// 1444 tap products (each rounded once, as in the sum) into LDS, and lane 0
// adds them in k_gabor_conv_t<false>'s (and the tile repair's) order: tap
// (j, i) reads input (y + 19 - j, x + 19 - i), +0 outside the image.
__global__ void __launch_bounds__(64) k_gabor_repair_list(const Img *imgs, const double *kernels, double *const *energy,
                                                          float *const *bound, const int *flags) {
  __shared__ double pre[GB_N * GB_N], pim[GB_N * GB_N];
  const int f = blockIdx.y;
  const int *fl = flags + (long)(blockIdx.z * 8 + f) * GR_STRIDE;
  const int cnt = fl[0];
  if (cnt == 0 || cnt > GR_CAP) return;
  const Img im = imgs[blockIdx.z];
  const long plane = (long)im.h * im.w;
  double *E = energy[blockIdx.z] + f * plane;
  float *D = bound[blockIdx.z] + f * plane;
  const gptr<const double> K = to_global(kernels + (long)f * GB_N * GB_N * 2);
  const gptr<const double> px = to_global((const double *)im.p);
  const int lane = threadIdx.x;
  for (int k = 0; k < cnt; k++) {
    const int i = fl[1 + k];
    const int y = i / im.w, x = i - y * im.w;
    for (int tp = lane; tp < GB_N * GB_N; tp += 64) {
      const int j = tp / GB_N, c = tp - j * GB_N;
      const int yy = y + 19 - j, xx = x + 19 - c;
      const double v = (yy >= 0 && yy < im.h && xx >= 0 && xx < im.w) ? px[(long)yy * im.w + xx] : 0.0;
      pre[tp] = v * K[2 * tp];
      pim[tp] = v * K[2 * tp + 1];
    }
    __syncthreads();
    if (lane == 0) {
      double re = 0.0, imv = 0.0;
      int tp = 0;
      for (; tp + 8 <= GB_N * GB_N; tp += 8) {  // eight products read at a time, added in order
        double a[8], b[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          a[u] = pre[tp + u];
          b[u] = pim[tp + u];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
          re += a[u];
          imv += b[u];
        }
      }
      for (; tp < GB_N * GB_N; tp++) {
        re += pre[tp];
        imv += pim[tp];
      }
      E[i] = sqrt(re * re + imv * imv);
      D[i] = 0.0f;
    }
    __syncthreads();  // the products of the next energy overwrite these
  }
}

__global__ void __launch_bounds__(256) k_gabor_repair(const Img *imgs, const double *kernels, double *const *energy,
                                                      float *const *bound, const int *flags, int loop_tiles) {
  __shared__ double tile[GB_TR + GB_N - 1][GB_LD];
  __shared__ double2 wk[GB_N * GB_N];
  __shared__ int list[GB_TR * GB_TC];
  __shared__ int nlist;
  const int f = blockIdx.y;
  if (flags[(long)(blockIdx.z * 8 + f) * GR_STRIDE] <= GR_CAP) return;  // none, or listed
  const Img im = imgs[blockIdx.z];
  const int ntiles = ((im.w + GB_TC - 1) / GB_TC) * ((im.h + GB_TR - 1) / GB_TR);
  const long plane = (long)im.h * im.w;
  double *E = energy[blockIdx.z] + f * plane;
  float *D = bound[blockIdx.z] + f * plane;
  const double *K = kernels + (long)f * GB_N * GB_N * 2;
  if (loop_tiles) {
    for (int tl = 0; tl < ntiles; tl++) gabor_repair_tile(im, tl, K, E, D, tile, wk, list, &nlist);
  } else if ((int)blockIdx.x < ntiles) {
    gabor_repair_tile(im, blockIdx.x, K, E, D, tile, wk, list, &nlist);
  }
}

__global__ void __launch_bounds__(1024) k_gabor_post_exact(int phase, const Img *imgs, double *const *energy,
                                                           float *const *bound, double *params, OutRef o,
                                                           int *flags) {
  __shared__ int hist[1024];
  __shared__ double shd[16];
  __shared__ long long shl[16];
  __shared__ int marks[8];            // energies marked per filter in this phase
  __shared__ int mlist[8][GR_CAP];    // the first GR_CAP of them
  if (threadIdx.x < 8) marks[threadIdx.x] = 0;
  __syncthreads();
  const Img im = imgs[blockIdx.x];
  const long n = (long)im.h * im.w;
  double *const E0 = energy[blockIdx.x];
  float *const D0 = bound[blockIdx.x];
  double *prm = params + (long)blockIdx.x * 8 * GP_STRIDE;
  const int f = blockIdx.y;  // one block per (image, filter)
  if (!(phase >= 4 && f == 0)) do {
    const double *E = E0 + (long)f * n;
    float *D = D0 + (long)f * n;
    double *P = prm + f * GP_STRIDE;
    auto mark = [&](long i) {
      D[i] = -D[i];
      const int k = atomicAdd(&marks[f], 1);
      if (k < GR_CAP) mlist[f][k] = (int)i;
    };
    if (phase == 1) {
      double lo = -1e308;
#pragma unroll 4
      for (long i = threadIdx.x; i < n; i += 1024) lo = fmax(lo, E[i] - (double)D[i]);
      lo = block_reduce(lo, MaxOp(), shd);
#pragma unroll 4
      for (long i = threadIdx.x; i < n; i += 1024)
        if (D[i] != 0.0f && E[i] + (double)D[i] >= lo) mark(i);
      break;
    }
    if (phase == 2) {
      double mx = -1e308;  // the candidates are exact now; nothing else can reach them
#pragma unroll 4
      for (long i = threadIdx.x; i < n; i += 1024)
        if (D[i] == 0.0f) mx = fmax(mx, E[i]);
      mx = block_reduce(mx, MaxOp(), shd);
      if (threadIdx.x == 0) P[0] = mx;
      if (!(mx > 0)) {  // no energy: every value exact, then the FP64 rules as written
#pragma unroll 4
        for (long i = threadIdx.x; i < n; i += 1024)
          if (D[i] != 0.0f) mark(i);
        break;
      }
      if (f == 0) {
        const double lvl = mx * 0.4;
#pragma unroll 4
        for (long i = threadIdx.x; i < n; i += 1024) {
          const double d = D[i];
          if (d != 0.0 && E[i] - d <= lvl && E[i] + d > lvl) mark(i);
        }
        break;
      }
      double hi_min = 1e308;
#pragma unroll 4
      for (long i = threadIdx.x; i < n; i += 1024) {
        const double v = (E[i] + (double)D[i]) / mx;
        if (moments_ok(v)) hi_min = fmin(hi_min, v);
      }
      hi_min = block_reduce(hi_min, MinOp(), shd);
#pragma unroll 4
      for (long i = threadIdx.x; i < n; i += 1024)
        if (D[i] != 0.0f && (E[i] - (double)D[i]) / mx <= hi_min) mark(i);
      break;
    }
    const double mx = P[0];
    if (phase == 3) {
      if (f == 0) {
        const double lvl = mx * 0.4;
        long long c = 0;
#pragma unroll 4
        for (long i = threadIdx.x; i < n; i += 1024) {
          const double d = D[i];
          c += d == 0.0 ? E[i] > lvl : E[i] - d > lvl;  // open pixels were recomputed
        }
        c = block_reduce(c, SumOp(), shl);
        if (threadIdx.x == 0) P[3] = (double)c;
        break;
      }
      double a = 1e308, b = -1e308;
#pragma unroll 4
      for (long i = threadIdx.x; i < n; i += 1024) {
        const double v = E[i] / mx;
        if (D[i] == 0.0f && moments_ok(v)) { a = fmin(a, v); b = fmax(b, v); }
      }
      a = block_reduce(a, MinOp(), shd);
      b = block_reduce(b, MaxOp(), shd);
      if (threadIdx.x == 0) { P[1] = a; P[2] = b; }
      const double sc = hist_scale(a, b, 1024);
#pragma unroll 4
      for (long i = threadIdx.x; i < n; i += 1024) {
        const double d = D[i];
        if (d == 0.0) continue;
        // the true value lies in [a, b]: clamp the interval's ends there
        const double vl = fmax((E[i] - d) / mx, a), vh = fmin((E[i] + d) / mx, b);
        if (hist_bin(vl, a, sc, 1024) != hist_bin(vh, a, sc, 1024)) mark(i);
      }
      break;
    }
    const double a = P[1], b = P[2];
    if (phase == 4) {
      for (int k = threadIdx.x; k < 1024; k += 1024) hist[k] = 0;
      __syncthreads();
      const double sc = hist_scale(a, b, 1024);
#pragma unroll 4
      for (long i = threadIdx.x; i < n; i += 1024) {
        const double e = E[i], d = D[i];
        // d > 0: the bin was decided by the interval (both ends give it)
        const double v = d == 0.0 ? e / mx : fmax((e - d) / mx, a);
        atomicAdd(&hist[hist_bin(v, a, sc, 1024)], 1);
      }
      __syncthreads();
      const double thr = otsu_threshold(hist, n, a, b);
      if (threadIdx.x == 0) P[3] = thr;
#pragma unroll 4
      for (long i = threadIdx.x; i < n; i += 1024) {
        const double d = D[i];
        if (d != 0.0 && (E[i] - d) / mx <= thr && (E[i] + d) / mx > thr) mark(i);
      }
      __syncthreads();  // hist is reused by the next filter
      break;
    }
    // phase 5
    const double thr = P[3];
    long long c = 0;
#pragma unroll 4
    for (long i = threadIdx.x; i < n; i += 1024) {
      const double d = D[i];
      c += d == 0.0 ? (E[i] / mx) > thr : (E[i] - d) / mx > thr;
    }
    c = block_reduce(c, SumOp(), shl);
    if (threadIdx.x == 0) o.rows[blockIdx.x][f - 1] = (double)c / prm[3];
  } while (0);
  __syncthreads();
  int *fl = flags + (long)(blockIdx.x * 8 + f) * GR_STRIDE;
  for (int k = threadIdx.x; k < GR_STRIDE; k += blockDim.x) {
    if (k == 0) fl[0] = marks[f];
    else if (k - 1 < marks[f] && k - 1 < GR_CAP) fl[k] = mlist[f][k - 1];
  }
}

void launch_gabor_post_exact(const Img *imgs, double *const *energy, float *const *bound, const double *taps,
                             double *params, int *flags, bool loop_tiles, OutRef o, int nimg, int maxh, int maxw,
                             hipStream_t s) {
  const int tiles = loop_tiles ? 1 : ((maxw + GB_TC - 1) / GB_TC) * ((maxh + GB_TR - 1) / GB_TR);
  for (int phase = 1; phase <= 5; phase++) {
    k_gabor_post_exact<<<dim3(nimg, 8), 1024, 0, s>>>(phase, imgs, energy, bound, params, o, flags);
    if (phase < 5) {
      k_gabor_repair_list<<<dim3(1, 8, nimg), 64, 0, s>>>(imgs, taps, energy, bound, flags);
      k_gabor_repair<<<dim3(tiles, 8, nimg), 256, 0, s>>>(imgs, taps, energy, bound, flags, loop_tiles);
    }
  }
}

void launch_gabor(const Img *imgs, const double *kernels, double *const *energy, OutRef o, int nimg, int maxh,
                  int maxw, hipStream_t s) {
  const int tiles = ((maxw + GB_TC - 1) / GB_TC) * ((maxh + GB_TR - 1) / GB_TR);
  k_gabor_conv_t<false><<<dim3(tiles, 8, nimg), 256, 0, s>>>(imgs, kernels, energy, nullptr, nullptr, 0.0);
  k_gabor_post<<<nimg, 1024, 0, s>>>(imgs, energy, o);
}
void launch_gabor_certified(const Img *imgs, const double *kernels, const double *wsum, double slack,
                            double *const *energy, float *const *bound, double *params, int *flags, OutRef o,
                            int nimg, int maxh, int maxw, hipStream_t s) {
  const int tiles = ((maxw + GB_TC - 1) / GB_TC) * ((maxh + GB_TR - 1) / GB_TR);
  k_gabor_conv_t<true><<<dim3(tiles, 8, nimg), 256, 0, s>>>(imgs, kernels, energy, wsum, bound, 4e-13 * slack);
  launch_gabor_post_exact(imgs, energy, bound, kernels, params, flags, slack <= 1.0, o, nimg, maxh, maxw, s);
}
void launch_gabor_post(const Img *imgs, double *const *energy, OutRef o, int nimg, hipStream_t s) {
  k_gabor_post<<<nimg, 1024, 0, s>>>(imgs, energy, o);
}

}  // namespace wnd
