// Adjoint (reverse-mode) kernels of the smooth transforms and features and
// of the WND similarity: the differentiable form of the feature path.
//
// Every kernel works on a batch of n planes of one size, FP64 row-major
// [n][h][w], with plain device pointers (the caller owns all buffers).  The
// forward values follow the feature path's definitions; the backward kernels
// recompute what they need from the input plane instead of keeping
// intermediate results (complex spectra, moment sums) between the passes.
#include "common.hpp"
#include "kernels.hpp"
#include "fft.hpp"

namespace wnd {

// ============================================================ Fourier |F|
// ImageMatrix::fft2 (src/cmatrix.cpp:1107-1160) is the unnormalised 2-D DFT
// magnitude, |F(ky,kx)| for ky < H/2+1 and the lower rows copied from
// |F(H-ky, (W-kx) mod W)|.  As a function of the pixels every output is
// |F| at its own frequency (Hermitian symmetry of a real plane), so
//   dL/dI(y,x) = Re sum_k Z(k) exp(-2 pi i (kx x / W + ky y / H)),
//   Z(k) = G(k) conj(F(k)) / |F(k)|   (0 where F(k) = 0),
// one more forward 2-D DFT of the complex plane Z.

// row spectra of real planes: buf[b][y][kx] = DFT_x(in[b][y][:])
__global__ void __launch_bounds__(256) k_ad_fft_rows_real(const double *in, double2 *buf, int h, int w,
                                                          FftPlan pl) {
  extern __shared__ double2 sh[];
  const long row = (long)blockIdx.y * h + blockIdx.x;
  const double *src = in + row * w;
  for (int x = threadIdx.x; x < w; x += blockDim.x) sh[x] = make_double2(src[x], 0.0);
  __syncthreads();
  double2 *r;
  fft_lds(sh, sh + w, pl, &r);
  double2 *o = buf + row * w;
  for (int x = threadIdx.x; x < w; x += blockDim.x) o[x] = r[x];
}

// column spectra -> magnitudes, upper rows computed and lower rows written
// from them exactly as the reference's fill
__global__ void __launch_bounds__(256) k_ad_fft_cols_mag(const double2 *buf, double *out, int h, int w,
                                                         FftPlan pl) {
  extern __shared__ double2 sh[];
  const int x = blockIdx.x;
  const double2 *src = buf + (long)blockIdx.y * h * w;
  for (int y = threadIdx.x; y < h; y += blockDim.x) sh[y] = src[(long)y * w + x];
  __syncthreads();
  double2 *r;
  fft_lds(sh, sh + h, pl, &r);
  double *o = out + (long)blockIdx.y * h * w;
  const int hh = h / 2 + 1, xm = x == 0 ? 0 : w - x;
  for (int y = threadIdx.x; y < hh; y += blockDim.x) {
    const double2 c = r[y];
    const double m = sqrt(c.x * c.x + c.y * c.y);
    o[(long)y * w + x] = m;
    if (y > 0 && h - y >= hh) o[(long)(h - y) * w + xm] = m;
  }
}

// backward, columns: F = DFT_y(column of the row spectra), Z = G conj(F)/|F|,
// then DFT_y(Z) written back over the same column
__global__ void __launch_bounds__(256) k_ad_fft_cols_adj(double2 *buf, const double *g, int h, int w, FftPlan pl) {
  extern __shared__ double2 sh[];
  const int x = blockIdx.x;
  double2 *col = buf + (long)blockIdx.y * h * w;
  const double *gp = g + (long)blockIdx.y * h * w;
  for (int y = threadIdx.x; y < h; y += blockDim.x) sh[y] = col[(long)y * w + x];
  __syncthreads();
  double2 *r;
  fft_lds(sh, sh + h, pl, &r);
  for (int y = threadIdx.x; y < h; y += blockDim.x) {
    const double2 c = r[y];
    const double m = sqrt(c.x * c.x + c.y * c.y);
    const double s = m > 0 ? gp[(long)y * w + x] / m : 0.0;
    r[y] = make_double2(s * c.x, -s * c.y);
  }
  __syncthreads();
  double2 *r2;
  fft_lds(r, r == sh ? sh + h : sh, pl, &r2);
  for (int y = threadIdx.x; y < h; y += blockDim.x) col[(long)y * w + x] = r2[y];
}

// backward, rows: grad[b][y][x] = Re DFT_x(buf[b][y][:])
__global__ void __launch_bounds__(256) k_ad_fft_rows_re(const double2 *buf, double *grad, int h, int w,
                                                        FftPlan pl) {
  extern __shared__ double2 sh[];
  const long row = (long)blockIdx.y * h + blockIdx.x;
  const double2 *src = buf + row * w;
  for (int x = threadIdx.x; x < w; x += blockDim.x) sh[x] = src[x];
  __syncthreads();
  double2 *r;
  fft_lds(sh, sh + w, pl, &r);
  double *o = grad + row * w;
  for (int x = threadIdx.x; x < w; x += blockDim.x) o[x] = r[x].x;
}

// lines up to 4096 points keep two complex buffers of 64 KiB each in LDS
static void fft_lds_attr() {
  static const bool done = [] {
    const int lds = 2 * 4096 * (int)sizeof(double2);
    (void)hipFuncSetAttribute((const void *)k_ad_fft_rows_real, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_ad_fft_cols_mag, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_ad_fft_cols_adj, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_ad_fft_rows_re, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    return true;
  }();
  (void)done;
}

void launch_ad_fourier_fwd(const double *in, double *out, double2 *buf, int n, int h, int w, const FftPlan &pw,
                           const FftPlan &ph, hipStream_t s) {
  fft_lds_attr();
  k_ad_fft_rows_real<<<dim3(h, n), 256, 2 * (size_t)w * sizeof(double2), s>>>(in, buf, h, w, pw);
  k_ad_fft_cols_mag<<<dim3(w, n), 256, 2 * (size_t)h * sizeof(double2), s>>>(buf, out, h, w, ph);
}

void launch_ad_fourier_bwd(const double *in, const double *g, double *grad, double2 *buf, int n, int h, int w,
                           const FftPlan &pw, const FftPlan &ph, hipStream_t s) {
  fft_lds_attr();
  k_ad_fft_rows_real<<<dim3(h, n), 256, 2 * (size_t)w * sizeof(double2), s>>>(in, buf, h, w, pw);
  k_ad_fft_cols_adj<<<dim3(w, n), 256, 2 * (size_t)h * sizeof(double2), s>>>(buf, g, h, w, ph);
  k_ad_fft_rows_re<<<dim3(h, n), 256, 2 * (size_t)w * sizeof(double2), s>>>(buf, grad, h, w, pw);
}

// ============================================================ strided GEMM
// C[b][r][c] = alpha * sum_a A[b](r, a) * B[b](a, c), C row-major M x N per
// batch entry.  The Chebyshev transform (src/cmatrix.cpp:1203-1222,
// chebyshev.cpp:118-165) is two such products against the basis tables and
// its adjoint is the two transposed ones; the Chebyshev-Fourier Downsample
// is two more.  FP64 on the matrix cores: a 64 x 64 tile per block, K staged
// through LDS 16 at a time, each wave a 32 x 32 sub-tile of 2 x 2
// v_mfma_f64_16x16x4 accumulators (operands: lane l holds A[l%16][k0+l/16]
// and B[k0+l/16][l%16]; result: register q of lane l holds row l/16 + 4q of
// column l%16).  Per 16-deep stage a wave reads 16 doubles per lane from LDS for
// 16 MFMAs (16K FMAs), against 8 reads per 16 FMAs for a register tile.
typedef double ad_d4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) k_ad_gemm(AdGemm d) {
  const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  const double *A = d.A + (long)blockIdx.z * d.sab;
  const double *B = d.B + (long)blockIdx.z * d.sbb;
  double *C = d.C + (long)blockIdx.z * d.scb;
  __shared__ double As[16][66];
  __shared__ double Bs[16][66];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int wr = (wv >> 1) * 32, wc = (wv & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  ad_d4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = ad_d4{0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < d.K; k0 += 16) {
    for (int e = threadIdx.x; e < 16 * 64; e += 256) {
      const int kk = e >> 6, rr = e & 63;
      const int k = k0 + kk, r = r0 + rr, c = c0 + rr;
      As[kk][rr] = (k < d.K && r < d.M) ? A[(long)r * d.sar + (long)k * d.sak] : 0.0;
      Bs[kk][rr] = (k < d.K && c < d.N) ? B[(long)k * d.sbk + (long)c * d.sbc] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; kk += 4) {
      const double a0 = As[kk + lk][wr + li], a1 = As[kk + lk][wr + 16 + li];
      const double b0 = Bs[kk + lk][wc + li], b1 = Bs[kk + lk][wc + 16 + li];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int r = r0 + wr + 16 * i + lk + 4 * q, c = c0 + wc + 16 * j + li;
        if (r < d.M && c < d.N) C[(long)r * d.N + c] = d.alpha * acc[i][j][q];
      }
}

void launch_ad_gemm(const AdGemm &d, int nbatch, hipStream_t s) {
  k_ad_gemm<<<dim3((d.N + 63) / 64, (d.M + 63) / 64, nbatch), 256, 0, s>>>(d);
}

// ============================================================ Zernike
// mb_zernike2D (src/textures/zernike/zernike.cpp:207-341): with the centroid
// (xc, yc) = (sum (i+1) f, sum (j+1) f) / S, S = sum f, and
// x = (i+1-xc)/rad, y = (j+1-yc)/rad, rad = min(W,H), each pixel of the
// closed unit disk (r >= DBL_EPSILON) adds
//   A_nm += (n+1)/pi * f/S * R_nm(r) * exp(-i (m+1) theta)
// (the reference's COST[m] / SINT[m] tables hold the angle (m+1) theta), and
// the feature is |A_nm|.  Writing R_nm(r) = r^m P_nm(r^2) and
// exp(-i q theta) = (x - i y)^q / r^q, q = m+1, the basis is
//   Z_nm(x, y) = c_n P_nm(t) (x - i y)^q / r,  t = r^2,  c_n = (n+1)/pi,
// a closed form whose x / y derivatives the adjoint needs:
//   dZ/dx = c_n [Q'(t) 2x w^q + Q(t) q w^(q-1)],
//   dZ/dy = c_n [Q'(t) 2y w^q - i Q(t) q w^(q-1)],
// with w = x - i y, Q = P / r, Q' = (P' - P / (2t)) / r.
// Because the centroid and S depend on every pixel,
//   dA/df_p = [Z(u_p) - A - x_p Gx - y_p Gy] / S,  Gx,y = sum_q f_q dZ/dx,y(u_q) / S.
// Pairs (n, m) are ordered n ascending, m ascending, n - m even (the
// feature order, zernike.cpp:329-337); pair index = zoff(n) + m/2.
constexpr int ZAD_L = 15;
constexpr int ZAD_PAIRS = 72;
constexpr int ZAD_NCO = 8;  // polynomial coefficients per pair (degree (n-m)/2 <= 7 in t)

__host__ __device__ constexpr int zad_off(int n) { return (n / 2) * (n / 2 + 1) + ((n & 1) ? n / 2 + 1 : 0); }

// P_nm coefficients, highest power of t first: c_s = (-1)^s (n-s)! /
// (s! ((n+m)/2-s)! ((n-m)/2-s)!), s = 0..(n-m)/2 (0 beyond), built at
// compile time into constant memory: every wave of a block reads the same
// pair at the same time, so the reads are scalar loads.
struct ZadTable {
  double co[ZAD_PAIRS * ZAD_NCO];
};
constexpr double zad_fact(int a) {
  double r = 1;
  for (int i = 2; i <= a; i++) r *= i;
  return r;
}
constexpr ZadTable zad_make() {
  ZadTable t{};
  for (int n = 0; n <= ZAD_L; n++)
    for (int m = n & 1; m <= n; m += 2) {
      const int k = zad_off(n) + m / 2, K = (n - m) / 2;
      for (int s = 0; s <= K; s++) {
        const double v = zad_fact(n - s) / (zad_fact(s) * zad_fact((n + m) / 2 - s) * zad_fact(K - s));
        t.co[k * ZAD_NCO + s] = (s & 1) ? -v : v;
      }
    }
  return t;
}
__constant__ ZadTable ZAD_TAB = zad_make();

// per image: hdr[b] = {S, xc, yc, rad}
__global__ void __launch_bounds__(1024) k_ad_zern_moments(const double *in, int h, int w, double *hdr) {
  __shared__ double shd[16];
  const double *p = in + (long)blockIdx.x * h * w;
  const long np = (long)h * w;
  double s = 0, m10 = 0, m01 = 0;
  for (long i = threadIdx.x; i < np; i += blockDim.x) {
    const int y = i / w, x = i - (long)y * w;
    const double v = p[i];
    s += v;
    m10 += (x + 1) * v;
    m01 += (y + 1) * v;
  }
  s = block_reduce(s, SumOp(), shd);
  m10 = block_reduce(m10, SumOp(), shd);
  m01 = block_reduce(m01, SumOp(), shd);
  if (threadIdx.x == 0) {
    double *o = hdr + 4 * (long)blockIdx.x;
    o[0] = s;
    o[1] = m10 / s;
    o[2] = m01 / s;
    o[3] = (double)(w < h ? w : h);
  }
}

// Moment sums over one pixel chunk for one group of nine pairs: the pairs
// of m = g and of m = 15 - g (8-g/2.. pairs each side, nine in all).
// part[((b * nchunk + chunk) * 72 + pair) * 6 + {Re A, Im A, Re Gx, Im Gx,
// Re Gy, Im Gy}] (unscaled sums of f * Z and f * dZ).
template <bool DERIV>
__global__ void __launch_bounds__(256) k_ad_zern_accum(const double *in, int h, int w, const double *hdr,
                                                       double *part, int nchunk) {
  __shared__ double red[4][9 * 6];
  const double *co = ZAD_TAB.co;
  const int b = blockIdx.z, g = blockIdx.y, chunk = blockIdx.x;
  const double *p = in + (long)b * h * w;
  const double S = hdr[4 * b], xc = hdr[4 * b + 1], yc = hdr[4 * b + 2], rad = hdr[4 * b + 3];
  const long np = (long)h * w, csz = (np + nchunk - 1) / nchunk;
  const long p0 = chunk * csz, p1 = p0 + csz < np ? p0 + csz : np;
  const int m1 = g, m2 = ZAD_L - g, cnt1 = (ZAD_L - m1) / 2 + 1;
  constexpr int NV = DERIV ? 6 : 2;
  double acc[9][NV];
#pragma unroll
  for (int k = 0; k < 9; k++)
#pragma unroll
    for (int v = 0; v < NV; v++) acc[k][v] = 0;
  const double PI = 3.14159265358979323846264338328;
  for (long q = p0 + threadIdx.x; q < p1; q += blockDim.x) {
    const int j = q / w, i = q - (long)j * w;
    const double x = (i + 1 - xc) / rad, y = (j + 1 - yc) / rad;
    const double t = x * x + y * y, r = sqrt(t);
    if (r < kDblEps || r > 1.0) continue;
    const double f = p[q], ir = 1.0 / r;
    // w^(q-1) and w^q for q = m1+1 and q = m2+1
    double a1r = 1, a1i = 0, a2r = 1, a2i = 0;
    for (int e = 0; e < m1; e++) {
      const double tr = a1r * x + a1i * y, ti = a1i * x - a1r * y;
      a1r = tr; a1i = ti;
    }
    for (int e = 0; e < m2; e++) {
      const double tr = a2r * x + a2i * y, ti = a2i * x - a2r * y;
      a2r = tr; a2i = ti;
    }
    const double b1r = a1r * x + a1i * y, b1i = a1i * x - a1r * y;
    const double b2r = a2r * x + a2i * y, b2i = a2i * x - a2r * y;
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const bool first = k < cnt1;
      const int m = first ? m1 : m2;
      const int n = first ? m1 + 2 * k : m2 + 2 * (k - cnt1);
      const int pk = zad_off(n) + (m >> 1);
      const int K = (n - m) >> 1;
      const double *c = co + pk * ZAD_NCO;
      double P = c[0], dP = 0;
      for (int s = 1; s <= K; s++) {
        dP = dP * t + P;
        P = P * t + c[s];
      }
      const double cn = (n + 1) / PI;
      const double wr = first ? b1r : b2r, wi = first ? b1i : b2i;   // w^q
      const double Q = P * ir;
      acc[k][0] += f * cn * Q * wr;
      acc[k][1] += f * cn * Q * wi;
      if constexpr (DERIV) {
        const double vr = first ? a1r : a2r, vi = first ? a1i : a2i;  // w^(q-1)
        const double dQ = (dP - P / (2 * t)) * ir;
        const double qq = (double)(m + 1) * Q;
        const double ex = 2 * x * dQ, ey = 2 * y * dQ;
        acc[k][2] += f * cn * (ex * wr + qq * vr);
        acc[k][3] += f * cn * (ex * wi + qq * vi);
        acc[k][4] += f * cn * (ey * wr + qq * vi);
        acc[k][5] += f * cn * (ey * wi - qq * vr);
      }
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 9; k++)
#pragma unroll
    for (int v = 0; v < NV; v++) {
      double s = acc[k][v];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
      if (lane == 0) red[wv][k * NV + v] = s;
    }
  __syncthreads();
  if (threadIdx.x < 9 * NV) {
    const int k = threadIdx.x / NV, v = threadIdx.x - k * NV;
    const double s = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
    const bool first = k < cnt1;
    const int m = first ? m1 : m2;
    const int n = first ? m1 + 2 * k : m2 + 2 * (k - cnt1);
    const int pk = zad_off(n) + (m >> 1);
    part[(((long)b * nchunk + chunk) * ZAD_PAIRS + pk) * 6 + v] = s;
  }
}

// Per image, one thread per pair: sums of the chunk partials in chunk
// order.  Forward: out[b][k] = |A_k|.  Backward (gout != null):
// coef[b] = {B_k = g_k conj(A_k)/|A_k| (72 complex), K0 = sum g|A|,
// Kx = Re sum B Gx, Ky = Re sum B Gy}.
__global__ void __launch_bounds__(128) k_ad_zern_final(const double *hdr, const double *part, int nchunk,
                                                       double *out, const double *gout, double *coef) {
  __shared__ double shd[16];
  const int b = blockIdx.x, k = threadIdx.x;
  const double S = hdr[4 * b];
  double e[6] = {0, 0, 0, 0, 0, 0};
  if (k < ZAD_PAIRS)
    for (int c = 0; c < nchunk; c++) {
      const double *pp = part + (((long)b * nchunk + c) * ZAD_PAIRS + k) * 6;
#pragma unroll
      for (int v = 0; v < 6; v++) e[v] += pp[v];
    }
  const double Ar = e[0] / S, Ai = e[1] / S;
  const double mag = sqrt(Ar * Ar + Ai * Ai);
  if (!gout) {
    if (k < ZAD_PAIRS) out[(long)b * ZAD_PAIRS + k] = mag;
    return;
  }
  double k0 = 0, kx = 0, ky = 0;
  if (k < ZAD_PAIRS) {
    const double g = gout[(long)b * ZAD_PAIRS + k];
    const double wgt = mag > 0 ? g / mag : 0.0;
    const double Br = wgt * Ar, Bi = -wgt * Ai;
    coef[(long)b * (2 * ZAD_PAIRS + 3) + 2 * k] = Br;
    coef[(long)b * (2 * ZAD_PAIRS + 3) + 2 * k + 1] = Bi;
    k0 = wgt * mag * mag;
    kx = Br * (e[2] / S) - Bi * (e[3] / S);
    ky = Br * (e[4] / S) - Bi * (e[5] / S);
  }
  k0 = block_reduce(k0, SumOp(), shd);
  kx = block_reduce(kx, SumOp(), shd);
  ky = block_reduce(ky, SumOp(), shd);
  if (k == 0) {
    double *cf = coef + (long)b * (2 * ZAD_PAIRS + 3) + 2 * ZAD_PAIRS;
    cf[0] = k0;
    cf[1] = kx;
    cf[2] = ky;
  }
}

// grad[b][p] = (sum_k Re(B_k Z_k(u_p)) - K0 - x_p Kx - y_p Ky) / S
__global__ void __launch_bounds__(256) k_ad_zern_pixels(int h, int w, const double *hdr, const double *coef,
                                                        double *grad) {
  __shared__ double B[2 * ZAD_PAIRS + 3];
  const double *co = ZAD_TAB.co;
  const int b = blockIdx.y;
  for (int i = threadIdx.x; i < 2 * ZAD_PAIRS + 3; i += blockDim.x) B[i] = coef[(long)b * (2 * ZAD_PAIRS + 3) + i];
  __syncthreads();
  const double S = hdr[4 * b], xc = hdr[4 * b + 1], yc = hdr[4 * b + 2], rad = hdr[4 * b + 3];
  const double K0 = B[2 * ZAD_PAIRS], Kx = B[2 * ZAD_PAIRS + 1], Ky = B[2 * ZAD_PAIRS + 2];
  const long np = (long)h * w;
  const double PI = 3.14159265358979323846264338328;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < np; q += (long)gridDim.x * blockDim.x) {
    const int j = q / w, i = q - (long)j * w;
    const double x = (i + 1 - xc) / rad, y = (j + 1 - yc) / rad;
    const double t = x * x + y * y, r = sqrt(t);
    double acc = 0;
    if (!(r < kDblEps || r > 1.0)) {
      const double ir = 1.0 / r;
      double wr = x, wi = -y;  // w^q, q = m + 1
      for (int m = 0; m <= ZAD_L; m++) {
        for (int n = m; n <= ZAD_L; n += 2) {
          const int pk = zad_off(n) + (m >> 1);
          const int K = (n - m) >> 1;
          const double *c = co + pk * ZAD_NCO;
          double P = c[0];
          for (int s = 1; s <= K; s++) P = P * t + c[s];
          acc += (n + 1) / PI * P * ir * (B[2 * pk] * wr - B[2 * pk + 1] * wi);
        }
        const double tr = wr * x + wi * y, ti = wi * x - wr * y;
        wr = tr; wi = ti;
      }
    }
    grad[(long)b * np + q] = (acc - K0 - x * Kx - y * Ky) / S;
  }
}

int zernike_ad_chunks(int h, int w) {
  const long np = (long)h * w;
  int c = (int)((np + 16383) / 16384);
  return c < 1 ? 1 : (c > 64 ? 64 : c);
}

void launch_ad_zernike_fwd(const double *in, double *out, double *work, int n, int h, int w, hipStream_t s) {
  const int nchunk = zernike_ad_chunks(h, w);
  double *hdr = work, *part = work + 4 * (long)n;
  k_ad_zern_moments<<<n, 1024, 0, s>>>(in, h, w, hdr);
  k_ad_zern_accum<false><<<dim3(nchunk, 8, n), 256, 0, s>>>(in, h, w, hdr, part, nchunk);
  k_ad_zern_final<<<n, 128, 0, s>>>(hdr, part, nchunk, out, nullptr, nullptr);
}

void launch_ad_zernike_bwd(const double *in, const double *gout, double *grad, double *work, int n, int h, int w,
                           hipStream_t s) {
  const int nchunk = zernike_ad_chunks(h, w);
  double *hdr = work, *part = work + 4 * (long)n;
  double *coef = part + (long)n * nchunk * ZAD_PAIRS * 6;
  k_ad_zern_moments<<<n, 1024, 0, s>>>(in, h, w, hdr);
  k_ad_zern_accum<true><<<dim3(nchunk, 8, n), 256, 0, s>>>(in, h, w, hdr, part, nchunk);
  k_ad_zern_final<<<n, 128, 0, s>>>(hdr, part, nchunk, nullptr, gout, coef);
  const long np = (long)h * w;
  int gx = (int)((np + 255) / 256);
  if (gx > 1024) gx = 1024;
  k_ad_zern_pixels<<<dim3(gx, n), 256, 0, s>>>(h, w, hdr, coef, grad);
}

size_t zernike_ad_work_doubles(int n, int h, int w) {
  return 4 * (size_t)n + (size_t)n * zernike_ad_chunks(h, w) * ZAD_PAIRS * 6 + (size_t)n * (2 * ZAD_PAIRS + 3);
}

// ============================================================ Chebyshev-Fourier
// ChebyshevFourier2D (src/transforms/ChebyshevFourier.cpp:59-160) over the
// (downsampled) plane v: with the in-disk pixels p and the host-built tables
// tn[p][a] = T_a(2r-1) (T_0 = 1/2), fc[p][i] = Ftrm cos(mf phi),
// fs[p][i] = Ftrm sin(-mf phi),
//   C[i][a] = sum_p v_p (fc[p][i] + i fs[p][i]) tn[p][a],  out = |C| (NN x NN),
// and the adjoint  dL/dv_p = sum_i fc[p][i] (sum_a Wr[i][a] tn[p][a])
//                          + fs[p][i] (sum_a Wi[i][a] tn[p][a]),  W = G C / |C|.
constexpr int CF_NN = 23;
constexpr int CF_STAGE = 32;  // pixels staged in LDS per step

// part[((b * nchunk + chunk) * 2 + {re, im}) * NN*NN + i*NN + a]
__global__ void __launch_bounds__(256) k_ad_cf_accum(const double *v, long plane, const int *kk, const double *tn,
                                                     const double *fc, const double *fs, int np, int NN, int nchunk,
                                                     double *part) {
  __shared__ double sv[CF_STAGE], stn[CF_STAGE][CF_NN], sfc[CF_STAGE][CF_NN], sfs[CF_STAGE][CF_NN];
  const int b = blockIdx.y, chunk = blockIdx.x;
  const double *img = v + (long)b * plane;
  const long csz = ((long)np + nchunk - 1) / nchunk;
  const long p0 = chunk * csz, p1 = p0 + csz < np ? p0 + csz : np;
  const int nout = NN * NN;
  double ar[3] = {0, 0, 0}, ai[3] = {0, 0, 0};
  int oi[3], oa[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int o = threadIdx.x + 256 * k;
    oi[k] = o < nout ? o / NN : 0;
    oa[k] = o < nout ? o - (o / NN) * NN : 0;
  }
  for (long q0 = p0; q0 < p1; q0 += CF_STAGE) {
    const int cnt = p1 - q0 < CF_STAGE ? (int)(p1 - q0) : CF_STAGE;
    __syncthreads();
    for (int e = threadIdx.x; e < cnt * NN; e += blockDim.x) {
      const int j = e / NN, a = e - j * NN;
      const long q = q0 + j;
      stn[j][a] = tn[q * NN + a];
      sfc[j][a] = fc[q * NN + a];
      sfs[j][a] = fs[q * NN + a];
    }
    for (int j = threadIdx.x; j < cnt; j += blockDim.x) sv[j] = img[kk[q0 + j]];
    __syncthreads();
    for (int j = 0; j < cnt; j++) {
      const double x = sv[j];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        ar[k] += (sfc[j][oi[k]] * stn[j][oa[k]]) * x;
        ai[k] += (sfs[j][oi[k]] * stn[j][oa[k]]) * x;
      }
    }
  }
  double *dst = part + ((long)b * nchunk + chunk) * 2 * nout;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int o = threadIdx.x + 256 * k;
    if (o < nout) {
      dst[o] = ar[k];
      dst[nout + o] = ai[k];
    }
  }
}

// per image: sums of the chunk partials in chunk order; forward writes |C|,
// backward (gout != null) writes W = G C / |C| (re, im) to coef[b][2][NN*NN]
__global__ void __launch_bounds__(256) k_ad_cf_final(const double *part, int nchunk, int NN, double *out,
                                                     const double *gout, double *coef) {
  const int b = blockIdx.x, nout = NN * NN;
  for (int o = threadIdx.x; o < nout; o += blockDim.x) {
    double cr = 0, ci = 0;
    for (int c = 0; c < nchunk; c++) {
      const double *pp = part + ((long)b * nchunk + c) * 2 * nout;
      cr += pp[o];
      ci += pp[nout + o];
    }
    const double mag = sqrt(cr * cr + ci * ci);
    if (!gout) {
      out[(long)b * nout + o] = mag;
    } else {
      const double s = mag > 0 ? gout[(long)b * nout + o] / mag : 0.0;
      coef[(long)b * 2 * nout + o] = s * cr;
      coef[(long)b * 2 * nout + nout + o] = s * ci;
    }
  }
}

// grad plane of the (downsampled) image: in-disk pixels get the adjoint,
// the rest 0 (one thread per plane pixel; `slot` maps a pixel to its
// in-disk index or -1)
__global__ void __launch_bounds__(256) k_ad_cf_pixels(const double *coef, const int *slot, const double *tn,
                                                      const double *fc, const double *fs, long plane, int NN,
                                                      double *grad) {
  __shared__ double W[2 * CF_NN * CF_NN];
  const int b = blockIdx.y, nout = NN * NN;
  for (int o = threadIdx.x; o < 2 * nout; o += blockDim.x) W[o] = coef[(long)b * 2 * nout + o];
  __syncthreads();
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < plane; p += (long)gridDim.x * blockDim.x) {
    const int q = slot[p];
    double acc = 0;
    if (q >= 0) {
      const double *t = tn + (long)q * NN, *c = fc + (long)q * NN, *sn = fs + (long)q * NN;
      double tv[CF_NN];
#pragma unroll
      for (int a = 0; a < CF_NN; a++) tv[a] = a < NN ? t[a] : 0.0;
      for (int i = 0; i < NN; i++) {
        double sr = 0, si = 0;
#pragma unroll
        for (int a = 0; a < CF_NN; a++) {
          if (a < NN) {
            sr += W[i * NN + a] * tv[a];
            si += W[nout + i * NN + a] * tv[a];
          }
        }
        acc += c[i] * sr + sn[i] * si;
      }
    }
    grad[(long)b * plane + p] = acc;
  }
}

__global__ void k_ad_fill_int(int *p, long n, int v) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_ad_cf_scatter(const int *kk, int np, int *slot) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < np) slot[kk[q]] = q;
}

// device fill (used where a memset would otherwise sit inside a capturable launch sequence)
void launch_fill_int(int *p, long n, int v, hipStream_t s) {
  if (n > 0) k_ad_fill_int<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, n, v);
}

// slot[pixel] = in-disk index or -1 (kk holds distinct indices below plane)
void launch_ad_cf_slots(const int *kk, int np, long plane, int *slot, hipStream_t s) {
  k_ad_fill_int<<<(unsigned)((plane + 255) / 256), 256, 0, s>>>(slot, plane, -1);
  k_ad_cf_scatter<<<(np + 255) / 256, 256, 0, s>>>(kk, np, slot);
}

int cf_ad_chunks(int np) {
  int c = (np + 2047) / 2048;
  return c < 1 ? 1 : c;
}

void launch_ad_cf_fwd(const double *v, int n, long plane, const int *kk, const double *tn, const double *fc,
                      const double *fs, int np, int NN, int nchunk, double *part, double *out, hipStream_t s) {
  k_ad_cf_accum<<<dim3(nchunk, n), 256, 0, s>>>(v, plane, kk, tn, fc, fs, np, NN, nchunk, part);
  k_ad_cf_final<<<n, 256, 0, s>>>(part, nchunk, NN, out, nullptr, nullptr);
}

void launch_ad_cf_bwd(const double *v, const double *gout, int n, long plane, const int *kk, const int *slot,
                      const double *tn, const double *fc, const double *fs, int np, int NN, int nchunk,
                      double *part, double *coef, double *grad, hipStream_t s) {
  k_ad_cf_accum<<<dim3(nchunk, n), 256, 0, s>>>(v, plane, kk, tn, fc, fs, np, NN, nchunk, part);
  k_ad_cf_final<<<n, 256, 0, s>>>(part, nchunk, NN, nullptr, gout, coef);
  long gx = (plane + 255) / 256;
  if (gx > 1024) gx = 1024;
  k_ad_cf_pixels<<<dim3((unsigned)gx, n), 256, 0, s>>>(coef, slot, tn, fc, fs, plane, NN, grad);
}

// ============================================================ Comb moments
// CombFirst4Moments2D (src/statistics/CombFirst4Moments.cpp:78-217): 4 combs
// x 20 positions ii = round(mp), mp = 1-m, 1-m+step, ... (step = m/10, at
// least 1, accumulated as the reference does); each line's Moments4 (mean,
// std for n > 2, skew = sqrt(n) M3 / M2^1.5 for n > 3, kurt = n M4 / M2^2
// for n > 4; src/statistics/Moments.h).  With I = y+1, J = x+1 the line
// conditions are integer equations:
//   comb 0: x - y = ii;  comb 1: x + y = n-1-ii;  comb 2: x = n/2-1-ii;  comb 3: y = m/2-1-ii
// (integer halves), so the lines of one comb are disjoint and the adjoint
// of a pixel sums one cubic in (v - mean) per comb:
//   dL/dv = c0 + c1 d + c2 d^2 + c3 d^3,  c1 = 2 G2, c2 = 3 G3, c3 = 4 G4,
//   c0 = g_mean / n - 3 G3 M2 / n - 4 G4 M3 / n,
// G2, G3, G4 the upstream gradient carried to M2, M3, M4.
constexpr int COMB_POS = 20;

__device__ int comb_positions(int m, int *ii) {
  double step = (double)m / 10.0;
  if (step < 1.0) step = 1.0;
  int c = 0;
  for (double mp = 1.0 - m; mp <= m; mp += step) {
    if (c < COMB_POS) ii[c] = (int)round(mp);
    c++;
  }
  return c < COMB_POS ? c : COMB_POS;
}

// pixel count and t-th pixel (row-major index) of line (comb, ii)
__device__ __forceinline__ int comb_line_len(int comb, int ii, int m, int n, int *first) {
  int y0, y1;
  switch (comb) {
    case 0: y0 = ii < 0 ? -ii : 0; y1 = (m - 1 < n - 1 - ii) ? m - 1 : n - 1 - ii; *first = y0; return y1 >= y0 ? y1 - y0 + 1 : 0;
    case 1: {
      const int sm = n - 1 - ii;
      y0 = sm - (n - 1) > 0 ? sm - (n - 1) : 0; y1 = (m - 1 < sm) ? m - 1 : sm; *first = y0;
      return y1 >= y0 ? y1 - y0 + 1 : 0;
    }
    case 2: { const int x = n / 2 - 1 - ii; *first = 0; return (x >= 0 && x < n) ? m : 0; }
    default: { const int y = m / 2 - 1 - ii; *first = 0; return (y >= 0 && y < m) ? n : 0; }
  }
}
__device__ __forceinline__ long comb_line_pixel(int comb, int ii, int m, int n, int first, int t) {
  switch (comb) {
    case 0: { const int y = first + t; return (long)y * n + (y + ii); }
    case 1: { const int y = first + t; return (long)y * n + (n - 1 - ii - y); }
    case 2: return (long)t * n + (n / 2 - 1 - ii);
    default: return (long)(m / 2 - 1 - ii) * n + t;
  }
}

// one block per (line, image): out[b][(comb * 4 + moment) * 20 + pos];
// backward (gout != null): coef[b][comb][pos] = {c0, c1, c2, c3, mean}
__global__ void __launch_bounds__(256) k_ad_comb_lines(const double *in, int h, int w, double *out,
                                                       const double *gout, double *coef) {
  __shared__ double shd[16];
  __shared__ int pos[COMB_POS];
  __shared__ int npos;
  const int b = blockIdx.y, comb = blockIdx.x / COMB_POS, p = blockIdx.x - comb * COMB_POS;
  if (threadIdx.x == 0) npos = comb_positions(h, pos);
  __syncthreads();
  const double *img = in + (long)b * h * w;
  int first = 0, L = 0;
  const int ii = p < npos ? pos[p] : 0;
  if (p < npos) L = comb_line_len(comb, ii, h, w, &first);
  double s = 0;
  for (int t = threadIdx.x; t < L; t += blockDim.x) s += img[comb_line_pixel(comb, ii, h, w, first, t)];
  s = block_reduce(s, SumOp(), shd);
  const double mu = L > 0 ? s / L : 0.0;
  double s2 = 0, s3 = 0, s4 = 0;
  for (int t = threadIdx.x; t < L; t += blockDim.x) {
    const double d = img[comb_line_pixel(comb, ii, h, w, first, t)] - mu;
    const double d2 = d * d;
    s2 += d2;
    s3 += d2 * d;
    s4 += d2 * d2;
  }
  const double M2 = block_reduce(s2, SumOp(), shd);
  const double M3 = block_reduce(s3, SumOp(), shd);
  const double M4 = block_reduce(s4, SumOp(), shd);
  if (threadIdx.x != 0) return;
  const double std = L > 2 ? sqrt(M2 / (L - 1)) : 0.0;
  const double skew = L > 3 ? (sqrt((double)L) * M3) / pow(M2, 1.5) : 0.0;
  const double kurt = L > 4 ? (L * M4) / (M2 * M2) : 0.0;
  if (!gout) {
    double *o = out + (long)b * 320 + comb * 80 + p;
    o[0] = mu;
    o[20] = std;
    o[40] = skew;
    o[60] = kurt;
    return;
  }
  const double *g = gout + (long)b * 320 + comb * 80 + p;
  double G2 = 0, G3 = 0, G4 = 0;
  if (M2 > 0) {
    if (L > 2) G2 += g[20] / (2 * std * (L - 1));
    if (L > 3) {
      G2 += g[40] * (-1.5 * sqrt((double)L) * M3 / pow(M2, 2.5));
      G3 = g[40] * sqrt((double)L) / pow(M2, 1.5);
    }
    if (L > 4) {
      G2 += g[60] * (-2.0 * L * M4 / (M2 * M2 * M2));
      G4 = g[60] * L / (M2 * M2);
    }
  }
  double *c = coef + ((long)b * 4 * COMB_POS + comb * COMB_POS + p) * 5;
  if (L > 0) {
    c[0] = g[0] / L - 3 * G3 * M2 / L - 4 * G4 * M3 / L;
    c[1] = 2 * G2;
    c[2] = 3 * G3;
    c[3] = 4 * G4;
  } else {
    c[0] = c[1] = c[2] = c[3] = 0;
  }
  c[4] = mu;
}

// grad[b][p] = sum over the four combs of the cubic of the pixel's line
__global__ void __launch_bounds__(256) k_ad_comb_pixels(const double *in, int h, int w, const double *coef,
                                                        double *grad) {
  __shared__ double cf[4 * COMB_POS * 5];
  __shared__ int pos[COMB_POS];
  __shared__ int npos;
  const int b = blockIdx.y;
  if (threadIdx.x == 0) npos = comb_positions(h, pos);
  for (int i = threadIdx.x; i < 4 * COMB_POS * 5; i += blockDim.x) cf[i] = coef[(long)b * 4 * COMB_POS * 5 + i];
  __syncthreads();
  const long np = (long)h * w;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < np; q += (long)gridDim.x * blockDim.x) {
    const int y = q / w, x = q - (long)y * w;
    const double v = in[(long)b * np + q];
    double acc = 0;
#pragma unroll
    for (int comb = 0; comb < 4; comb++) {
      const int ii = comb == 0 ? x - y : comb == 1 ? w - 1 - x - y : comb == 2 ? w / 2 - 1 - x : h / 2 - 1 - y;
      int k = -1;
      for (int j = 0; j < npos; j++)
        if (pos[j] == ii) k = j;
      if (k >= 0) {
        const double *c = cf + (comb * COMB_POS + k) * 5;
        const double d = v - c[4];
        acc += c[0] + d * (c[1] + d * (c[2] + d * c[3]));
      }
    }
    grad[(long)b * np + q] = acc;
  }
}

void launch_ad_comb_fwd(const double *in, double *out, int n, int h, int w, hipStream_t s) {
  k_ad_comb_lines<<<dim3(4 * COMB_POS, n), 256, 0, s>>>(in, h, w, out, nullptr, nullptr);
}

void launch_ad_comb_bwd(const double *in, const double *gout, double *grad, double *coef, int n, int h, int w,
                        hipStream_t s) {
  k_ad_comb_lines<<<dim3(4 * COMB_POS, n), 256, 0, s>>>(in, h, w, nullptr, gout, coef);
  const long np = (long)h * w;
  long gx = (np + 255) / 256;
  if (gx > 1024) gx = 1024;
  k_ad_comb_pixels<<<dim3((unsigned)gx, n), 256, 0, s>>>(in, h, w, coef, grad);
}

// ============================================================ Gabor energy
// GaborEnergy (src/textures/gabor.cpp:166-195): the full complex convolution
// of the plane with the 38 x 38 kernel G_f (conv2comp, 19-52) cropped at
// ceil(38/2) = 19, i.e. a correlation with the flipped kernel:
//   re + i im = sum_{ju, iv} G_f[37-ju][37-iv] I[b - 18 + ju][a - 18 + iv]  (I = 0 outside),
//   E_f(b, a) = |re + i im|,
// for f0 = 0.1 (low pass) then 1..7 (GaborTextureFilters2D, 201-231).  With
// upstream G_E and Z = G_E (re, im) / E the adjoint is a correlation with
// the kernel itself:
//   dL/dI(l, k) = sum_f sum_{u, v} Re Z_f(l - 19 + u, k - 19 + v) G_re[u][v] + Im Z_f(..) G_im[u][v].
// One block per 16-row output tile of one filter; the input halo in LDS,
// the taps read as uniform (broadcast) loads, a sliding register window
// along each tap row.
constexpr int GB_N = 38, GB_TH = 16;

// MODE 0: out = E (plane of the [n][8][h][w] maps); MODE 1: out = Z (complex [n][h][w])
template <int MODE>
__global__ void __launch_bounds__(256) k_ad_gabor_fwd(const double *in, const double *taps, int h, int w,
                                                      const double *gE, long estride, double *out) {
  constexpr int TW = 48, Q = 3;  // 16 x 48 outputs, three per thread along a row
  __shared__ double S[GB_TH + GB_N - 1][TW + GB_N - 1];
  const int b = blockIdx.z, y0 = blockIdx.y * GB_TH, x0 = blockIdx.x * TW;
  const double *img = in + (long)b * h * w;
  for (int e = threadIdx.x; e < (GB_TH + GB_N - 1) * (TW + GB_N - 1); e += blockDim.x) {
    const int r = e / (TW + GB_N - 1), c = e - r * (TW + GB_N - 1);
    const int y = y0 - 18 + r, x = x0 - 18 + c;
    S[r][c] = (y >= 0 && y < h && x >= 0 && x < w) ? img[(long)y * w + x] : 0.0;
  }
  __syncthreads();
  const int r = threadIdx.x >> 4, c = (threadIdx.x & 15) * Q;
  double re[Q] = {0, 0, 0}, im[Q] = {0, 0, 0};
  for (int ju = 0; ju < GB_N; ju++) {
    const double *trow = taps + (long)(GB_N - 1 - ju) * GB_N * 2;
    const double *srow = &S[r + ju][c];
    double w0 = srow[0], w1 = srow[1];
    for (int iv = 0; iv < GB_N; iv++) {
      const double w2 = srow[iv + 2];
      const double gr = trow[(GB_N - 1 - iv) * 2], gi = trow[(GB_N - 1 - iv) * 2 + 1];
      re[0] = fma(gr, w0, re[0]); im[0] = fma(gi, w0, im[0]);
      re[1] = fma(gr, w1, re[1]); im[1] = fma(gi, w1, im[1]);
      re[2] = fma(gr, w2, re[2]); im[2] = fma(gi, w2, im[2]);
      w0 = w1;
      w1 = w2;
    }
  }
  const int y = y0 + r;
  if (y >= h) return;
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const int x = x0 + c + q;
    if (x >= w) continue;
    const double e = sqrt(re[q] * re[q] + im[q] * im[q]);
    if (MODE == 0) {
      out[(long)b * estride + (long)y * w + x] = e;
    } else {
      const double g = gE[(long)b * estride + (long)y * w + x];
      const double sc = e > 0 ? g / e : 0.0;
      out[((long)b * h * w + (long)y * w + x) * 2] = sc * re[q];
      out[((long)b * h * w + (long)y * w + x) * 2 + 1] = sc * im[q];
    }
  }
}

// grad (+)= sum_{u,v} Re Z(l - 19 + u, k - 19 + v) G_re[u][v] + Im Z(..) G_im[u][v]
__global__ void __launch_bounds__(256) k_ad_gabor_adj(const double *Z, const double *taps, int h, int w, int acc,
                                                      double *grad) {
  constexpr int TW = 32, Q = 2;  // 16 x 32 outputs, two per thread
  __shared__ double2 S[GB_TH + GB_N - 1][TW + GB_N - 1];
  const int b = blockIdx.z, y0 = blockIdx.y * GB_TH, x0 = blockIdx.x * TW;
  const double2 *zp = (const double2 *)Z + (long)b * h * w;
  for (int e = threadIdx.x; e < (GB_TH + GB_N - 1) * (TW + GB_N - 1); e += blockDim.x) {
    const int r = e / (TW + GB_N - 1), c = e - r * (TW + GB_N - 1);
    const int y = y0 - 19 + r, x = x0 - 19 + c;
    S[r][c] = (y >= 0 && y < h && x >= 0 && x < w) ? zp[(long)y * w + x] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  const int r = threadIdx.x >> 4, c = (threadIdx.x & 15) * Q;
  double g[Q] = {0, 0};
  for (int u = 0; u < GB_N; u++) {
    const double *trow = taps + (long)u * GB_N * 2;
    const double2 *srow = &S[r + u][c];
    double2 w0 = srow[0];
    for (int v = 0; v < GB_N; v++) {
      const double2 w1 = srow[v + 1];
      const double gr = trow[v * 2], gi = trow[v * 2 + 1];
      g[0] = fma(gr, w0.x, fma(gi, w0.y, g[0]));
      g[1] = fma(gr, w1.x, fma(gi, w1.y, g[1]));
      w0 = w1;
    }
  }
  const int y = y0 + r;
  if (y >= h) return;
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const int x = x0 + c + q;
    if (x >= w) continue;
    double *o = grad + (long)b * h * w + (long)y * w + x;
    *o = acc ? *o + g[q] : g[q];
  }
}

// forward: the 8 maps into out [n][8][h][w]
void launch_ad_gabor_fwd(const double *in, const double *taps8, double *out, int n, int h, int w, hipStream_t s) {
  const dim3 grid((w + 47) / 48, (h + GB_TH - 1) / GB_TH, n);
  for (int f = 0; f < 8; f++)
    k_ad_gabor_fwd<0><<<grid, 256, 0, s>>>(in, taps8 + (long)f * GB_N * GB_N * 2, h, w, nullptr, 8L * h * w,
                                          out + (long)f * h * w);
}

// backward: per filter, Z from a recomputed convolution, then its correlation added into grad
void launch_ad_gabor_bwd(const double *in, const double *taps8, const double *gE, double *grad, double *Z, int n,
                         int h, int w, hipStream_t s) {
  const dim3 gf((w + 47) / 48, (h + GB_TH - 1) / GB_TH, n), ga((w + 31) / 32, (h + GB_TH - 1) / GB_TH, n);
  for (int f = 0; f < 8; f++) {
    const double *t = taps8 + (long)f * GB_N * GB_N * 2;
    k_ad_gabor_fwd<1><<<gf, 256, 0, s>>>(in, t, h, w, gE + (long)f * h * w, 8L * h * w, Z);
    k_ad_gabor_adj<<<ga, 256, 0, s>>>(Z, t, h, w, f > 0, grad);
  }
}

// ============================================================ WND similarity
// classify2 (src/wndchrm_src/TrainingSet.cpp:2429-2586) after
// signatures::normalize (signatures.cpp:196-215):
//   v_f = 100 (x_f - min_f) / (max_f - min_f) (0 / 100 outside, 0 for NaN),
//   d_t = sum_f w_f^2 (v_f - T_tf)^2 over terms with w_f, |v_f - T_tf| >= eps,
//   s_t = d_t^-5 (d_t >= eps), sbar_c = mean of s_t over class c, p_c = sbar_c / sum sbar.
// With upstream G = dL/dp:
//   dL/dsbar_c = (G_c - sum_k G_k p_k) / sum sbar,
//   dL/dd_t = -5 d_t^-6 dL/dsbar_c(t) / n_c(t),
//   dL/dv_f = sum_t dL/dd_t 2 w_f^2 (v_f - T_tf),  dL/dx_f = dL/dv_f * 100 / (max_f - min_f).

// per test vector: gsn[q][c] = dL/dsbar_c / n_c (0 for empty classes)
__global__ void k_ad_wnd_cls(const double *dist, const int *cls, int n, int C, int ntest, const double *prob,
                             const double *norm, const double *gp, double *gsn) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ntest) return;
  int nper[64];
  for (int c = 0; c <= C; c++) nper[c] = 0;
  for (int i = 0; i < n; i++) {
    const int ci = cls[i];
    if (!(dist[(long)q * n + i] < kDblEps) && ci >= 0 && ci <= C) nper[ci]++;
  }
  double dot = 0;
  for (int c = 0; c < C; c++) dot += gp[(long)q * C + c] * prob[(long)q * C + c];
  for (int c = 1; c <= C; c++) {
    const double gs = (gp[(long)q * C + c - 1] - dot) / norm[q];
    gsn[(long)q * (C + 1) + c] = nper[c] ? gs / nper[c] : 0.0;
  }
}

// gd[q][t] = dL/dd_t
__global__ void k_ad_wnd_gd(const double *dist, const int *cls, int n, int C, int ntest, const double *gsn,
                            double *gd) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)n * ntest) return;
  const int q = e / n, i = e - (long)q * n;
  const double d = dist[e];
  const int ci = cls[i];
  gd[e] = (d < kDblEps || ci < 1 || ci > C) ? 0.0 : -5.0 * pow(d, -6.0) * gsn[(long)q * (C + 1) + ci];
}

// gx[q][f]: one thread per (q, f), training rows in order
__global__ void k_ad_wnd_gx(const double *train, const double *wts, const double *x, const double *v,
                            const double *mins, const double *maxs, const double *gd, int n, int f, int ntest,
                            double *gx) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)f * ntest) return;
  const int q = e / f, s = e - (long)q * f;
  const double ws = wts[s];
  double acc = 0;
  if (!(ws < kDblEps)) {
    const double vs = v[e];
    for (int i = 0; i < n; i++) {
      const double dlt = vs - train[(long)i * f + s];
      if (fabs(dlt) < kDblEps) continue;
      acc += gd[(long)q * n + i] * dlt;
    }
    acc *= 2 * ws * ws;
  }
  const double xv = x[e], mn = mins[s], mx = maxs[s];
  const bool lin = !(xv != xv || xv < mn || (mx - mn) < kDblEps || xv > mx);
  gx[e] = lin ? acc * (100.0 / (mx - mn)) : 0.0;
}

void launch_ad_wnd_bwd(const double *train, const int *cls, const double *wts, const double *x, const double *v,
                       const double *mins, const double *maxs, const double *dist, const double *prob,
                       const double *norm, const double *gp, double *gsn, double *gd, int n, int f, int C, int ntest,
                       double *gx, hipStream_t s) {
  k_ad_wnd_cls<<<(ntest + 63) / 64, 64, 0, s>>>(dist, cls, n, C, ntest, prob, norm, gp, gsn);
  const long e = (long)n * ntest;
  k_ad_wnd_gd<<<(e + 255) / 256, 256, 0, s>>>(dist, cls, n, C, ntest, gsn, gd);
  const long e2 = (long)f * ntest;
  k_ad_wnd_gx<<<(e2 + 255) / 256, 256, 0, s>>>(train, wts, x, v, mins, maxs, gd, n, f, ntest, gx);
}

}  // namespace wnd
