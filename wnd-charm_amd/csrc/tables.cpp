// Host-side geometry tables.  Every value that depends only on image
// geometry (never on pixel data) is computed here with the host C library,
// using the reference's own expressions, and uploaded once per size.  This
// keeps transcendental-function rounding identical to the reference's.
#include "tables.hpp"

#include <cmath>
#include <cstring>

namespace wnd {

// Chebyshev basis tj[a][j] (src/transforms/chebyshev.cpp:38-78 TNx, 80-101 getChCoeff1D):
// x_a = 2(a+1)/len - 1; T = cos(j*acos(x)) (T0 = 1, |x| > 1 -> acos arg 0);
// tj = T/len for j = 0, T*2/len otherwise.
std::vector<double> cheb_basis(int len, int N) {
  std::vector<double> x(len), T((size_t)len * N), B((size_t)len * N);
  for (int a = 0; a < len; a++) x[a] = 2 * (double)(a + 1) / (double)len - 1;
  for (int iy = 0; iy < len; iy++)
    for (int ix = 0; ix < N; ix++) {
      double t = x[iy];
      t = (fabs(t) > 1) ? 0 : acos(t);
      double o = t * (double)ix;
      T[(size_t)iy * N + ix] = cos(o);
    }
  for (int iy = 0; iy < len; iy++) T[(size_t)iy * N] = 1;
  for (int a = 0; a < len; a++)
    for (int j = 0; j < N; j++) {
      double tj = T[(size_t)a * N + j];
      B[(size_t)a * N + j] = j ? tj * 2 / (double)len : tj / (double)len;
    }
  return B;
}

// FFT twiddles exp(-2 pi i k / n) and radix list (smallest radices first)
void fft_tables(int n, std::vector<double> &tw, std::vector<int> &radix) {
  tw.resize(2 * (size_t)n);
  for (int k = 0; k < n; k++) {
    const double a = 2.0 * M_PI * (double)k / (double)n;
    tw[2 * k] = cos(a);
    tw[2 * k + 1] = -sin(a);
  }
  radix.clear();
  int r = n;
  for (int f : {4, 2, 3, 5, 7})
    while (r % f == 0 && r > 1) { radix.push_back(f); r /= f; }
  for (int f = 11; f * f <= r; f += 2)
    while (r % f == 0) { radix.push_back(f); r /= f; }
  if (r > 1) radix.push_back(r);
  if (radix.empty()) radix.push_back(1);
}

// Chebyshev-Fourier geometry (src/transforms/ChebyshevFourier.cpp:42-128)
CfTable cf_table(int w, int h) {
  CfTable t;
  const unsigned long m = h, n = w;
  const double two_n = (2.0 / ((double)n - 1)), two_m = (2.0 / ((double)m - 1));
  unsigned long N = 11;
  const unsigned long Nmax = (unsigned long)(((m < n ? m : n) - 1) / 2);
  if (N > Nmax) N = Nmax;
  const unsigned long NN = 2 * N + 1;
  t.NN = (int)NN;
  std::vector<double> Tn(NN);
  for (unsigned long x = 0; x < n; x++) {
    const double x_ind = -1.0 + (double)x * two_n;
    const double x_2 = x_ind * x_ind;
    for (unsigned long y = 0; y < m; y++) {
      const double y_ind = -1.0 + (double)y * two_m;
      const double r = sqrt(x_2 + y_ind * y_ind);
      const double f = -1 * atan2(y_ind, x_ind);
      if (!(r < 1)) continue;
      t.kk.push_back((int)(y * n + x));
      // ChebPol(2r - 1, NN)
      double xx = r * 2 - 1;
      if (xx < -1) xx = -1;
      if (xx > 1) xx = 1;
      const double acos_x = acos(xx);
      for (unsigned long a = 0; a < NN; a++) Tn[a] = cos(a * acos_x);
      Tn[0] = 0.5;
      for (unsigned long a = 0; a < NN; a++) t.tn.push_back(Tn[a]);
      // Ftrm*cos(mf*phi), Ftrm*sin(-mf*phi) for mf = -N..N.  (-k)*phi is
      // -(k*phi) exactly and libm's cos is even and sin odd bit for bit, so
      // each |mf| takes one cos and one sin call.
      double ck[12], sk[12];
      for (unsigned long k = 1; k <= N; k++) {
        ck[k] = cos((double)k * f);
        sk[k] = sin((double)k * f);
      }
      for (unsigned long im = 1; im <= NN; im++) {
        const long mf = (long)im - 1 - (long)N;
        if (mf == 0) {
          t.fc.push_back(0.5 * cos(mf * f));
          t.fs.push_back(0.5 * sin(-1 * mf * f));
        } else {
          const unsigned long k = (unsigned long)(mf < 0 ? -mf : mf);
          t.fc.push_back(1.0 * ck[k]);
          t.fs.push_back(1.0 * (mf < 0 ? sk[k] : -sk[k]));
        }
      }
    }
  }
  return t;
}

// step positions of ImageMatrix::Downsample (cmatrix.cpp:579-720):
// x_k by the reference's sequential x += dx while x < len
std::vector<double> ds_positions(unsigned len, double dx) {
  std::vector<double> xs;
  double x = 0;
  while (x < len) {
    xs.push_back(x);
    x += dx;
  }
  return xs;
}

// One pass of Downsample as a dense [nout][len] matrix: row k holds the
// weights ds_line gives input a for output k (fractional ends, sequential
// positions), divided by dx.  The differentiable form applies it as a GEMM
// and its transpose.
std::vector<double> ds_dense(unsigned len, double dx, unsigned nout) {
  const std::vector<double> xs = ds_positions(len, dx);
  std::vector<double> D((size_t)nout * len, 0.0);
  for (unsigned k = 0; k < nout && k < xs.size(); k++) {
    const double x = xs[k];
    double *row = &D[(size_t)k * len];
    unsigned a = (unsigned)(floor(x));
    double frac = ceil(x) - x;
    if (frac > 0 && a < len) row[a] += frac;
    for (a = (unsigned)(ceil(x)); a < floor(x + dx); a++)
      if (a < len) row[a] += 1;
    frac = x + dx - floor(x + dx);
    if (frac > 0 && a < len) row[a] += frac;
    for (unsigned i = 0; i < len; i++) row[i] /= dx;
  }
  return D;
}

// Gabor filters (gabor.cpp:103-162), theta = 3.14159265/2, gamma 0.5,
// sig2lam 0.56, n = 38; index 0: f0 = 0.1, 1..7: f0 = 1..7
std::vector<double> gabor_kernels() {
  const double gamma = 0.5, sig2lam = 0.56, theta = 3.14159265 / 2, fi = 0;
  const int n = 38;
  std::vector<double> all;
  const double f0s[8] = {0.1, 1, 2, 3, 4, 5, 6, 7};
  for (int fidx = 0; fidx < 8; fidx++) {
    const double f0 = f0s[fidx];
    const double lambda = 2 * M_PI / f0;
    const double ct = cos(theta), st = sin(theta);
    const double sig = sig2lam * lambda;
    std::vector<double> tx(n + 1), G((size_t)n * n * 2);
    tx[0] = -(n / 2);
    for (int x = 1; x <= n; x++) tx[x] = tx[x - 1] + 1;
    const std::vector<double> ty(tx);
    double sum = 0;
    for (int y = 0; y < n; y++)
      for (int x = 0; x < n; x++) {
        const double xte = tx[x] * ct + ty[y] * st;
        const double yte = ty[y] * ct - tx[x] * st;
        const double rte = xte * xte + gamma * gamma * yte * yte;
        const double ge = exp(-1 * rte / (2 * sig * sig));
        const double argm = xte * f0 + fi;
        G[y * n * 2 + x * 2] = ge * cos(argm);
        G[y * n * 2 + x * 2 + 1] = ge * sin(argm);
        sum += sqrt(G[y * n * 2 + x * 2] * G[y * n * 2 + x * 2] + G[y * n * 2 + x * 2 + 1] * G[y * n * 2 + x * 2 + 1]);
      }
    for (int y = 0; y < n; y++)
      for (int x = 0; x < n * 2; x++) G[y * n * 2 + x] = G[y * n * 2 + x] / sum;
    all.insert(all.end(), G.begin(), G.end());
  }
  return all;
}

// Zernike H1/H2/H3 recurrence coefficients (zernike.cpp:253-264), [3][32][32]
std::vector<double> zernike_h() {
  std::vector<double> h(3 * 32 * 32, 0.0);
  double *H1 = h.data(), *H2 = H1 + 1024, *H3 = H2 + 1024;
  for (int n = 0; n < 32; n++)
    for (int m = 0; m <= n; m++)
      if (n != m) {
        H3[n * 32 + m] = -(double)(4.0 * (m + 2.0) * (m + 1.0)) / (double)((n + m + 2.0) * (n - m));
        H2[n * 32 + m] = ((double)(H3[n * 32 + m] * (n + m + 4.0) * (n - m - 2.0)) / (double)(4.0 * (m + 3.0))) + (m + 2.0);
        H1[n * 32 + m] = ((double)((m + 4.0) * (m + 3.0)) / 2.0) - ((m + 4.0) * H2[n * 32 + m]) +
                         ((double)(H3[n * 32 + m] * (n + m + 6.0) * (n - m - 4.0)) / 8.0);
      }
  return h;
}

// Radon angle table (radon.cpp:52-56): cos / sin of theta*M_PI/180
std::vector<double> radon_trig() {
  std::vector<double> t(8);
  const double th[4] = {0, 45, 90, 135};
  for (int k = 0; k < 4; k++) {
    const double angle = (th[k] * M_PI) / 180;
    t[2 * k] = cos(angle);
    t[2 * k + 1] = sin(angle);
  }
  return t;
}

// Comb line offsets (CombFirst4Moments.cpp:113-117): ii = round(m_pos),
// m_pos = 1 - m, += m/10 (>= 1) while m_pos <= m; at most 20 used, NaN pads
std::vector<double> comb_offsets(int m) {
  std::vector<double> ii(20, std::nan(""));
  double step = (double)m / 10.0;
  if (step < 1.0) step = 1.0;
  int k = 0;
  for (double mp = 1.0 - m; mp <= m; mp += step) {
    if (k < 20) ii[k] = (double)(long)round(mp);
    k++;
  }
  return ii;
}

}  // namespace wnd
