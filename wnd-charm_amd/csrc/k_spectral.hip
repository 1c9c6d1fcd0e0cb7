// Moment / basis contractions: Zernike coefficients, Chebyshev-Fourier
// coefficients (with the reference's area-weighted downsampling for planes
// above 300x300 pixels).  FP64.
#include "common.hpp"
#include "kernels.hpp"

namespace wnd {

// ===================================================================== Zernike
// mb_zernike2D (src/textures/zernike/zernike.cpp:207-341), order 15, radius
// min(W,H), unit circle centred on the intensity centroid.  Per-pixel values
// (radial powers, angle-addition cos/sin, the H1/H2/H3 radial recurrence)
// use the reference's operations in the reference's order; the sums over
// pixels are a fixed-order tree.
constexpr int ZL = 15;
constexpr int ZPAIRS = 72;  // (n, m), 0 <= m <= n <= 15, n - m even
__host__ __device__ constexpr int zoff(int n) {  // pairs before order n
  int s = 0;
  for (int k = 0; k < n; k++) s += k / 2 + 1;
  return s;
}

// header per image in partial[]: [0] m10/m00, [1] m01/m00, [2] sum
__global__ void __launch_bounds__(1024) k_zern_moments(const Img *imgs, double *partial, int nchunk) {
  __shared__ double shd[16];
  const Img im = imgs[blockIdx.x];
  const long n = (long)im.h * im.w;
  double s = 0, m10 = 0, m01 = 0;
  for (long i = threadIdx.x; i < n; i += blockDim.x) {
    const int y = i / im.w, x = i - (long)y * im.w;
    const double v = im.p[i];
    s += v;
    m10 += (x + 1) * v;
    m01 += (y + 1) * v;
  }
  s = block_reduce(s, SumOp(), shd);
  m10 = block_reduce(m10, SumOp(), shd);
  m01 = block_reduce(m01, SumOp(), shd);
  if (threadIdx.x == 0) {
    double *hdr = partial + (long)blockIdx.x * (4 + (long)nchunk * 2 * ZPAIRS);
    hdr[0] = m10 / s;
    hdr[1] = m01 / s;
    hdr[2] = s;
  }
}

// One wavefront per group of orders [N0, N1] (groups balanced by work),
// lanes over pixels; the group's accumulators stay in registers (every loop
// is unrolled at compile time) and are reduced across lanes at the end.
// T: the accumulators' type (double for the plain form; float for the
// reduced-precision forms, where the radial polynomials stay FP64).
// Zernike coefficients are floating-point features (held to 1e-6 of the
// reference), so this function, unlike the exact kernels, lets the compiler
// contract multiply-adds into FMAs and replaces the reference's divisions
// by products with reciprocals: of the radius and of the plane's sum once
// per block, of r once per pixel (1/r^2 and cos, sin of the angle from it),
// and of pi once.
template <int N0, int N1, class T>
__device__ __forceinline__ void zern_group(const Img im, long p0, long p1, int lane, double cxm, double cym, double sum,
                                           const double *__restrict__ H1, const double *__restrict__ H2,
                                           const double *__restrict__ H3, double *__restrict__ dst) {
#pragma clang fp contract(fast)
  constexpr int NP = [] {
    int c = 0;
    for (int n = N0; n <= N1; n++) c += n / 2 + 1;
    return c;
  }();
  const double rad = (double)(im.w < im.h ? im.w : im.h);
  const double irad = 1.0 / rad, isum = 1.0 / sum;
  const double IPI = 1.0 / 3.14159265358979323846264338328;
  T AR[NP], AI[NP];
#pragma unroll
  for (int k = 0; k < NP; k++) AR[k] = AI[k] = (T)0;
  for (long p = p0 + lane; p < p1; p += 64) {
    const int j = p / im.w, i = p - (long)j * im.w;  // i: column, j: row
    const double x = (i + 1 - cxm) * irad;
    const double y = (j + 1 - cym) * irad;
    const double r2 = x * x + y * y;
    const double r = sqrt(r2);
    if (r < kDblEps || r > 1.0) continue;
    const double ir = 1.0 / r, ir2 = ir * ir;
    double R[N1 + 1];
    R[0] = 1;
#pragma unroll
    for (int k = 1; k <= N1; k++) R[k] = r * R[k - 1];
    double COST[N1 + 1], SINT[N1 + 1];
    const double a = COST[0] = x * ir;
    const double b = SINT[0] = y * ir;
#pragma unroll
    for (int m = 1; m <= N1; m++) {
      COST[m] = a * COST[m - 1] - b * SINT[m - 1];
      SINT[m] = a * SINT[m - 1] + b * COST[m - 1];
    }
    const double f = im.p[p] * isum;
    int base = 0;
#pragma unroll
    for (int n = N0; n <= N1; n++) {
      const double const_t = (n + 1) * f * IPI;
      const double Rn = R[n];
      const double Rnm2 = n >= 2 ? R[n >= 2 ? n - 2 : 0] : 0.0;
      double Rnmp2 = 0, Rnmp4 = 0, Rnm = 0;
#pragma unroll
      for (int m = n; m >= 0; m -= 2) {
        if (m == n) { Rnm = Rn; Rnmp4 = Rn; }
        else if (m == n - 2) { const double Rnnm2 = n * Rn - (n - 1) * Rnm2; Rnm = Rnnm2; Rnmp2 = Rnnm2; }
        else {
          Rnm = H1[n * 32 + m] * Rnmp4 + (H2[n * 32 + m] + H3[n * 32 + m] * ir2) * Rnmp2;
          Rnmp4 = Rnmp2;
          Rnmp2 = Rnm;
        }
        AR[base + (m >> 1)] += (T)(const_t * Rnm * COST[m]);
        AI[base + (m >> 1)] -= (T)(const_t * Rnm * SINT[m]);
      }
      base += n / 2 + 1;
    }
  }
  const int off = zoff(N0);
#pragma unroll
  for (int k = 0; k < NP; k++) {
    T vr = AR[k], vi = AI[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      vr += __shfl_xor(vr, o, 64);
      vi += __shfl_xor(vi, o, 64);
    }
    if (lane == 0) {
      dst[2 * (off + k)] = vr;
      dst[2 * (off + k) + 1] = vi;
    }
  }
}

// block = one pixel chunk of one image; wave w handles order group w
template <class T>
__global__ void __launch_bounds__(256) k_zern_accum(const Img *__restrict__ imgs, double *__restrict__ partial, int nchunk,
                                                    const double *__restrict__ htab) {
  const Img im = imgs[blockIdx.y];
  const long P = (long)im.h * im.w;
  const long csz = (P + nchunk - 1) / nchunk;
  const long p0 = blockIdx.x * csz, p1 = (p0 + csz < P) ? p0 + csz : P;
  double *hdr = partial + (long)blockIdx.y * (4 + (long)nchunk * 2 * ZPAIRS);
  const double cxm = hdr[0], cym = hdr[1], sum = hdr[2];
  const double *H1 = htab, *H2 = htab + 32 * 32, *H3 = htab + 2 * 32 * 32;
  double *dst = hdr + 4 + (long)blockIdx.x * 2 * ZPAIRS;
  const int lane = threadIdx.x & 63;
  switch (threadIdx.x >> 6) {
    case 0: zern_group<0, 7, T>(im, p0, p1, lane, cxm, cym, sum, H1, H2, H3, dst); break;
    case 1: zern_group<8, 10, T>(im, p0, p1, lane, cxm, cym, sum, H1, H2, H3, dst); break;
    case 2: zern_group<11, 13, T>(im, p0, p1, lane, cxm, cym, sum, H1, H2, H3, dst); break;
    default: zern_group<14, 15, T>(im, p0, p1, lane, cxm, cym, sum, H1, H2, H3, dst); break;
  }
}

__global__ void k_zern_final(const Img *imgs, const double *partial, int nchunk, OutRef o) {
  const int e = threadIdx.x;
  if (e >= ZPAIRS) return;
  const double *hdr = partial + (long)blockIdx.x * (4 + (long)nchunk * 2 * ZPAIRS);
  double ar = 0, ai = 0;
  for (int c = 0; c < nchunk; c++) {
    ar += hdr[4 + (long)c * 2 * ZPAIRS + 2 * e];
    ai += hdr[4 + (long)c * 2 * ZPAIRS + 2 * e + 1];
  }
  ar *= ar;
  ai *= ai;
  o.rows[blockIdx.x][e] = fabs(sqrt(ar + ai));
}

void launch_zernike(const Img *imgs, double *partial, int nchunk, const double *htab, OutRef o, int nimg,
                    hipStream_t s, int precision) {
  k_zern_moments<<<nimg, 1024, 0, s>>>(imgs, partial, nchunk);
  // FP64 accumulation in every precision: the radial polynomials are FP64
  // anyway, and without packed FP32 instructions (this library is built
  // with them off) FP32 and FP64 FMAs issue at the same rate, so FP32 sums
  // would add error (above 1e-5 on transformed planes) and save nothing
  (void)precision;
  k_zern_accum<double><<<dim3(nchunk, nimg), 256, 0, s>>>(imgs, partial, nchunk, htab);
  k_zern_final<<<nimg, 128, 0, s>>>(imgs, partial, nchunk, o);
}

// ============================================================ Chebyshev-Fourier
// ChebyshevFourier2D (src/transforms/ChebyshevFourier.cpp:59-160), N = 11,
// NN = 23: sum_r[im][a] += (Ftrm*cos(mf*phi) * T_a(2r-1)) * I  over in-disk
// pixels; the geometry tables are host-computed with the reference's libm
// calls.  The 529 magnitudes only feed a 32-bin histogram, so the sums are
// taken on the FP64 matrix cores and every bin is certified against an
// a-priori bound on the difference from the reference's sequential sums;
// an entry with a bin left open is recomputed in the reference's order.
constexpr int CF_CHUNK = 64;
constexpr int CF_NN = 23;

typedef double cf_d4 __attribute__((ext_vector_type(4)));

// One K-step (4 pixels) of NW entries.  The sums of an entry are the GEMM
//   S[r][a] = sum_p A[r][p] B[p][a],  A[r][p] = fc[p][r] (r < NN), fs[p][r-NN] (NN <= r < 2NN),
//   B[p][a] = tn[p][a] I[p],
// 3 row tiles x 2 column tiles of v_mfma_f64_16x16x4 (lane l holds
// A[l%16][k0+l/16] and B[k0+l/16][l%16]); rows >= 2NN and columns >= NN
// are zero in LDS.
template <int NW>
__device__ __forceinline__ void cfm_step(cf_d4 (&acc)[4][6], const double (*Acat)[48], const double (*Tcat)[32],
                                         const double (*iv)[CF_CHUNK], int e0, int s, int lane) {
  const int p = 4 * s + (lane >> 4), c = lane & 15;
  const double a0 = Acat[p][c], a1 = Acat[p][16 + c], a2 = Acat[p][32 + c];
  const double t0 = Tcat[p][c], t1 = Tcat[p][16 + c];
#pragma unroll
  for (int e = 0; e < NW; e++) {
    const double v = iv[e0 + e][p];
    const double b0 = t0 * v, b1 = t1 * v;
    acc[e][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[e][0], 0, 0, 0);
    acc[e][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[e][1], 0, 0, 0);
    acc[e][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[e][2], 0, 0, 0);
    acc[e][3] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[e][3], 0, 0, 0);
    acc[e][4] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, b0, acc[e][4], 0, 0, 0);
    acc[e][5] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, b1, acc[e][5], 0, 0, 0);
  }
}

template <int NW>
__device__ __forceinline__ void cfm_chunk(cf_d4 (&acc)[4][6], const double (*Acat)[48], const double (*Tcat)[32],
                                          const double (*iv)[CF_CHUNK], int e0, int lane) {
#pragma unroll 4
  for (int s = 0; s < CF_CHUNK / 4; s++) cfm_step<NW>(acc, Acat, Tcat, iv, e0, s, lane);
}

// Block (split x, group y) = CFM_GROUP entries sharing one table set; wave w
// owns entries 4w..4w+3 over the chunk range of split x.  The tables of a
// chunk are staged in LDS once for all 16 entries; the next chunk's table
// entries and pixels are fetched into registers while the current one is
// consumed.  Output per (entry, split): the 6 accumulator tiles as held in
// the MFMA registers ([tile][q][lane]) and sum |I| over the split's pixels.
__global__ void __launch_bounds__(256) k_cf_mfma(const CfMGroup *groups, int nsplit) {
  __shared__ double Acat[CF_CHUNK][48];
  __shared__ double Tcat[CF_CHUNK][32];
  __shared__ double iv[CFM_GROUP][CF_CHUNK];
  const CfMGroup &G = groups[blockIdx.y];
  const int NN = G.NN, ne = G.ne, np = G.np;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  for (int k = t; k < CF_CHUNK * 48; k += 256) (&Acat[0][0])[k] = 0.0;
  for (int k = t; k < CF_CHUNK * 32; k += 256) (&Tcat[0][0])[k] = 0.0;
  const int nchunk = (np + CF_CHUNK - 1) / CF_CHUNK;
  const int c0 = (int)((long)blockIdx.x * nchunk / nsplit), c1 = (int)((long)(blockIdx.x + 1) * nchunk / nsplit);
  constexpr int CF_PER = (CF_CHUNK * CF_NN + 255) / 256;
  const gptr<const double> Gtn = to_global(G.tn), Gfc = to_global(G.fc), Gfs = to_global(G.fs);
  double rt[CF_PER], rc[CF_PER], rs[CF_PER], rv[4];
  // staging roles: table element t + 256k; pixel t % 64 of entries t/64 + 4j
  const int q = t & 63;
  auto fetch = [&](int ch) {
    const int p0 = ch * CF_CHUNK;
    const int npc = min(np - p0, CF_CHUNK);
    const long base = (long)p0 * NN;
#pragma unroll
    for (int k = 0; k < CF_PER; k++) {
      const int e = min(t + 256 * k, npc * NN - 1);  // past the chunk: re-read its last entry (unused)
      rt[k] = Gtn[base + e];
      rc[k] = Gfc[base + e];
      rs[k] = Gfs[base + e];
    }
    const int kidx = to_global(G.kk)[p0 + min(q, npc - 1)];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int e = wv + 4 * j;
      rv[j] = (e < ne && q < npc) ? to_global(G.img[e])[kidx] : 0.0;
    }
  };
  cf_d4 acc[4][6];
#pragma unroll
  for (int e = 0; e < 4; e++)
#pragma unroll
    for (int k = 0; k < 6; k++) acc[e][k] = cf_d4{0.0, 0.0, 0.0, 0.0};
  double sabs[4] = {0.0, 0.0, 0.0, 0.0};
  const int nw = max(0, min(4, ne - 4 * wv));
  if (c0 < c1) fetch(c0);
  for (int ch = c0; ch < c1; ch++) {
    const int npc = min(np - ch * CF_CHUNK, CF_CHUNK);
    __syncthreads();  // the previous chunk is consumed
#pragma unroll
    for (int k = 0; k < CF_PER; k++) {
      const int e = t + 256 * k;
      if (e < npc * NN) {
        const int p = e / NN, a = e - p * NN;
        Tcat[p][a] = rt[k];
        Acat[p][a] = rc[k];
        Acat[p][NN + a] = rs[k];
      }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      iv[wv + 4 * j][q] = rv[j];  // zero past the chunk's pixels and for absent entries
      sabs[j] += fabs(rv[j]);
    }
    __syncthreads();
    if (ch + 1 < c1) fetch(ch + 1);
    switch (nw) {
      case 4: cfm_chunk<4>(acc, Acat, Tcat, iv, 4 * wv, lane); break;
      case 3: cfm_chunk<3>(acc, Acat, Tcat, iv, 4 * wv, lane); break;
      case 2: cfm_chunk<2>(acc, Acat, Tcat, iv, 4 * wv, lane); break;
      case 1: cfm_chunk<1>(acc, Acat, Tcat, iv, 4 * wv, lane); break;
      default: break;
    }
  }
#pragma unroll
  for (int e = 0; e < 4; e++) {
    if (e < nw) {
      const gptr<double> P = to_global(G.part[4 * wv + e] + (long)blockIdx.x * CFM_PART);
#pragma unroll
      for (int k = 0; k < 6; k++)
#pragma unroll
        for (int r = 0; r < 4; r++) P[k * 256 + r * 64 + lane] = acc[e][k][r];
    }
  }
  // sum |I| of entries wv + 4j over this split's pixels (a bound input only)
#pragma unroll
  for (int j = 0; j < 4; j++) {
    double v = sabs[j];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0 && wv + 4 * j < ne) to_global(G.part[wv + 4 * j] + (long)blockIdx.x * CFM_PART)[6 * 256] = v;
  }
}

// element (r, a) of the 48x32 padded sum matrix in the [tile][q][lane] layout
__device__ __forceinline__ int cfm_index(int r, int a) {
  const int tile = (r >> 4) * 2 + (a >> 4), rr = r & 15;
  return tile * 256 + (rr >> 2) * 64 + (rr & 3) * 16 + (a & 15);
}

// Certified histogram of one entry.  With |Ftrm cos|, |Ftrm sin| <= 1 and
// |T_a| <= 1, every term of a coefficient sum is at most |I_p|, so both the
// reference's sequential sum (K - 1 additions after two products per term)
// and the matrix-core sum (one product for B, the MFMA product and at most
// K + nsplit + 8 additions on any term's path) are within
// gamma_{K+nsplit+8} * sum|I| of the exact sum of the terms (Higham,
// Accuracy and Stability of Numerical Algorithms, 2nd ed., eq. 4.4), FP64
// subnormals adding at most 2^-1074 per operation.  Each magnitude then lies
// within e = 1.5 E + 8u c of the reference's (the 2-norm is 1-Lipschitz;
// pow, the sum and sqrt round within 4u on either side), and the
// reference's min, max and every (c - min) * 32 / (max - min) lie in
// intervals from those; a bin is certified when both ends of its interval,
// widened by 1e-12 relative for the rounding of the reference's and this
// kernel's own arithmetic, floor to the same bin.  open[e] = 1 marks an
// entry whose bins are not all certified (its row is then left unwritten).
__global__ void __launch_bounds__(256) k_cf_certify(const CfDesc *cds, int *open, double slack, OutRef o) {
  __shared__ double shd[16];
  __shared__ double cc[CF_NN * CF_NN], ce[CF_NN * CF_NN];
  __shared__ int bins[32];
  __shared__ int bad;
  const CfDesc &C = cds[blockIdx.x];
  const int NN = C.NN, NN2 = NN * NN, nsp = C.nsplit;
  const gptr<const double> P = to_global((const double *)C.part);
  if (threadIdx.x == 0) bad = 0;
  if (threadIdx.x < 32) bins[threadIdx.x] = 0;
  double S = 0;
  for (int s = 0; s < nsp; s++) S += P[(long)s * CFM_PART + 6 * 256];
  const double u = 0x1p-53;
  const double n = (double)C.np + nsp + 8;
  const double gam = n * u / (1 - n * u);
  const double E = slack * (gam * S * (1 + 1e-6) + (n + 16) * 0x1p-1070);
  double mnL = kInf, mnU = kInf, mxL = -kInf, mxU = -kInf;
  bool fin = isfinite(E) && n * u < 0.01;
  __syncthreads();
  for (int k = threadIdx.x; k < NN2; k += blockDim.x) {
    const int im = k / NN, a = k - im * NN;
    const int ir = cfm_index(im, a), ii = cfm_index(NN + im, a);
    double sr = 0, si = 0;
    for (int s = 0; s < nsp; s++) {
      sr += P[(long)s * CFM_PART + ir];
      si += P[(long)s * CFM_PART + ii];
    }
    const double c = sqrt(sr * sr + si * si);
    const double e = 2 * E + 16 * u * c + 0x1p-1000;
    fin = fin && isfinite(c) && isfinite(e);
    cc[k] = c;
    ce[k] = e;
    mnL = fmin(mnL, c - e);
    mnU = fmin(mnU, c + e);
    mxL = fmax(mxL, c - e);
    mxU = fmax(mxU, c + e);
  }
  if (!fin) bad = 1;
  mnL = block_reduce(mnL, MinOp(), shd);
  mnU = block_reduce(mnU, MinOp(), shd);
  mxL = block_reduce(mxL, MaxOp(), shd);
  mxU = block_reduce(mxU, MaxOp(), shd);
  const double den_lo = mxL - mnU, den_hi = mxU - mnL;
  if (!(den_lo > 0) || bad) {
    if (threadIdx.x == 0) open[blockIdx.x] = 1;
    return;
  }
  for (int k = threadIdx.x; k < NN2; k += blockDim.x) {
    const double c = cc[k], e = ce[k];
    const double num_lo = fmax(0.0, (c - e) - mnU), num_hi = (c + e) - mnL;
    const double tl = 32 * num_lo / den_hi * (1 - 1e-12);
    const double tu = 32 * num_hi / den_lo * (1 + 1e-12) + 1e-300;
    const int bl = tl >= 31 ? 31 : (int)tl, bu = tu >= 31 ? 31 : (int)tu;
    if (bl != bu) bad = 1;
    atomicAdd(&bins[bl], 1);
  }
  __syncthreads();
  if (bad) {
    if (threadIdx.x == 0) open[blockIdx.x] = 1;
    return;
  }
  if (threadIdx.x == 0) open[blockIdx.x] = 0;
  if (threadIdx.x < 32) o.rows[blockIdx.x][threadIdx.x] = bins[threadIdx.x];
}

// Reference-order recomputation of the groups holding an open entry.
// Block = one group of entries sharing the basis tables.  Each thread owns
// up to five of the 2*NN*NN coefficients and keeps one running sum per
// coefficient and entry across all pixel chunks, so every sum is taken in
// the reference's pixel order (ChebyshevFourier.cpp:112-135); the
// tables of a chunk are staged in LDS once for all entries of the group.
__global__ void __launch_bounds__(256) k_cf_accum(const CfGroup *groups, const int *open) {
  __shared__ double tn[CF_CHUNK][CF_NN], fc[CF_CHUNK][CF_NN], fs[CF_CHUNK][CF_NN];
  __shared__ double iv[CF_GROUP][CF_CHUNK];
  const CfGroup G = groups[blockIdx.x];
  const int NN = G.NN, NN2 = NN * NN, ne = G.ne;
  bool any = false;
  for (int g = 0; g < ne; g++) any = any || open[G.ent[g]];
  if (!any) return;
  const int t = threadIdx.x;
  // this thread's coefficients: o = t + 256 q, q < 5 (2*23*23 = 1058 <= 1280)
  int oim[5], oa[5];
  bool ore[5], on[5];
#pragma unroll
  for (int q = 0; q < 5; q++) {
    const int o = t + 256 * q;
    on[q] = o < 2 * NN2;
    const int part = o >= NN2, k = part ? o - NN2 : o;
    ore[q] = !part;
    oim[q] = on[q] ? k / NN : 0;
    oa[q] = on[q] ? k - (k / NN) * NN : 0;
  }
  double acc[CF_GROUP][5];
#pragma unroll
  for (int g = 0; g < CF_GROUP; g++)
#pragma unroll
    for (int q = 0; q < 5; q++) acc[g][q] = 0.0;
  constexpr int CF_PER = (CF_CHUNK * CF_NN + 255) / 256;
  const gptr<const double> Gtn = to_global(G.tn), Gfc = to_global(G.fc), Gfs = to_global(G.fs);
  double rt[CF_PER], rc[CF_PER], rs[CF_PER], rv[CF_GROUP];
  auto fetch = [&](int p0) {
    const int np = (G.np - p0 < CF_CHUNK) ? G.np - p0 : CF_CHUNK;
    const long base = (long)p0 * NN;
#pragma unroll
    for (int k = 0; k < CF_PER; k++) {
      const int e = min(t + 256 * k, np * NN - 1);  // past the chunk: re-read its last entry (unused)
      rt[k] = Gtn[base + e];
      rc[k] = Gfc[base + e];
      rs[k] = Gfs[base + e];
    }
    const int kidx = to_global(G.kk)[p0 + min(t, np - 1)];
#pragma unroll
    for (int g = 0; g < CF_GROUP; g++) rv[g] = g < ne ? to_global((const double *)G.img[g])[kidx] : 0.0;
  };
  fetch(0);
  for (int p0 = 0; p0 < G.np; p0 += CF_CHUNK) {
    const int np = (G.np - p0 < CF_CHUNK) ? G.np - p0 : CF_CHUNK;
    __syncthreads();  // the previous chunk is consumed
#pragma unroll
    for (int k = 0; k < CF_PER; k++) {
      const int e = t + 256 * k;
      if (e < np * NN) {
        const int p = e / NN, a = e - p * NN;
        tn[p][a] = rt[k];
        fc[p][a] = rc[k];
        fs[p][a] = rs[k];
      }
    }
    if (t < np) {
#pragma unroll
      for (int g = 0; g < CF_GROUP; g++) iv[g][t] = rv[g];
    }
    __syncthreads();
    if (p0 + CF_CHUNK < G.np) fetch(p0 + CF_CHUNK);
    for (int p = 0; p < np; p++) {
#pragma unroll
      for (int q = 0; q < 5; q++) {
        if (!on[q]) continue;
        const double f = ore[q] ? fc[p][oim[q]] : fs[p][oim[q]];
        const double ft = f * tn[p][oa[q]];
#pragma unroll
        for (int g = 0; g < CF_GROUP; g++) acc[g][q] = acc[g][q] + ft * iv[g][p];
      }
    }
  }
#pragma unroll
  for (int g = 0; g < CF_GROUP; g++)
    if (g < ne)
#pragma unroll
      for (int q = 0; q < 5; q++)
        if (on[q]) G.sum[g][t + 256 * q] = acc[g][q];
}

// Histogram of the reference-order sums of the open entries
// (ChebyshevFourier.cpp:137-153).
__global__ void __launch_bounds__(1024) k_cf_final(const CfDesc *cds, const int *open, OutRef o) {
  __shared__ double shd[16];
  __shared__ double coeff[CF_NN * CF_NN];
  __shared__ int bins[32];
  if (!open[blockIdx.x]) return;
  const CfDesc C = cds[blockIdx.x];
  const int NN = C.NN, NN2 = NN * NN;
  double mn = kInf, mx = -kInf;
  for (int k = threadIdx.x; k < NN2; k += blockDim.x) {
    const double sr = C.sum[k], si = C.sum[NN2 + k];
    const double v = sqrt(sr * sr + si * si);
    coeff[k] = v;
    mn = fmin(mn, v);
    mx = fmax(mx, v);
  }
  mn = block_reduce(mn, MinOp(), shd);
  mx = block_reduce(mx, MaxOp(), shd);
  if (threadIdx.x < 32) bins[threadIdx.x] = 0;
  __syncthreads();
  if (mx - mn > 0) {
    const double hs = (double)32 / (double)(mx - mn);
    for (int k = threadIdx.x; k < NN2; k += blockDim.x) atomicAdd(&bins[hist_bin(coeff[k], mn, hs, 32)], 1);
  }
  __syncthreads();
  if (threadIdx.x < 32) o.rows[blockIdx.x][threadIdx.x] = bins[threadIdx.x];
}

void launch_cf(const CfMGroup *mgroups, int nmgroups, int nsplit, const CfGroup *groups, int ngroups,
               const CfDesc *cd, int nimg, int *open, double slack, OutRef o, hipStream_t s) {
  k_cf_mfma<<<dim3(nsplit, nmgroups), 256, 0, s>>>(mgroups, nsplit);
  k_cf_certify<<<nimg, 256, 0, s>>>(cd, open, slack, o);
  k_cf_accum<<<ngroups, 256, 0, s>>>(groups, open);
  k_cf_final<<<nimg, 1024, 0, s>>>(cd, open, o);
}

// ================================================================== Downsample
// ImageMatrix::Downsample (cmatrix.cpp:579-720), grayscale: each output is
// the reference's left-fraction + whole pixels + right-fraction sum / dx, with
// the step positions xs[k] (ys[k]) precomputed by the same sequential
// x += dx accumulation.
__device__ __forceinline__ double ds_line(const double *v, long stride, unsigned len, double x, double dx) {
  double sum_i = 0;
  unsigned a = (unsigned)(floor(x));
  double frac = ceil(x) - x;
  if (frac > 0 && a < len) sum_i += v[(long)a * stride] * frac;
  for (a = (unsigned)(ceil(x)); a < floor(x + dx); a++)
    if (a < len) sum_i += v[(long)a * stride];
  frac = x + dx - floor(x + dx);
  if (frac > 0 && a < len) sum_i += v[(long)a * stride] * frac;
  return sum_i;
}

__global__ void k_ds_x(const DsDesc *dds) {
  const DsDesc D = dds[blockIdx.y];
  const long n = (long)D.oh * D.cx;
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int y = t / D.cx, k = t - (long)y * D.cx;
  D.tmp[(long)y * D.ow + k] = ds_line(D.in + (long)y * D.ow, 1, D.ow, D.xs[k], D.dx) / (D.dx);
}

__global__ void k_ds_y(const DsDesc *dds) {
  const DsDesc D = dds[blockIdx.y];
  const long n = (long)D.nh * D.nw;
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int y = t / D.nw, x = t - (long)y * D.nw;
  D.out[(long)y * D.nw + x] = ds_line(D.tmp + x, D.ow, D.oh, D.ys[y], D.dy) / D.dy;
}

void launch_downsample(const DsDesc *dd, int nimg, int maxoh, int maxow, int maxnh, int maxnw, hipStream_t s) {
  const long n1 = (long)maxoh * maxow, n2 = (long)maxnh * maxnw;
  k_ds_x<<<dim3((n1 + 255) / 256, nimg), 256, 0, s>>>(dd);
  k_ds_y<<<dim3((n2 + 255) / 256, nimg), 256, 0, s>>>(dd);
}

}  // namespace wnd
