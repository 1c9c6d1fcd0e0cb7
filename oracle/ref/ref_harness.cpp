/*
 * ref_harness.cpp -- C entry points over reference functions that compile
 * from their own sources (no generated config.h, no FFTW, no ImageMatrix
 * out-of-line members).  TEST INFRASTRUCTURE: used only to pin the oracle.
 *
 *   ref_texture  -> Extract_Texture_Features (src/textures/haralick/CVIPtexture.cpp:384-452)
 *   ref_eigen    -> Eigen 3.1.2 eigenvalues() of a row-major matrix, the call
 *                   f14_maxcorr makes on Q.matrix() (CVIPtexture.cpp:1015)
 *   ref_radon    -> radon + vd_RadonTextures input path (src/transforms/radon.cpp:14-91)
 *   ref_wavelet  -> Symlet5(0,1).transform2D on a DataGrid2D
 *                   (src/transforms/wavelet/Symlet5.cpp:10-59, Wavelet.cpp:136-434)
 *   ref_cheb1d   -> TNx + getChCoeff (src/transforms/chebyshev.cpp:38-116)
 */
#include <cstring>
#include <vector>
#include <Eigen/Core>
#include <Eigen/Eigenvalues>
#include "textures/haralick/CVIPtexture.h"
#include "transforms/radon.h"
#include "colors/FuzzyCalc.h"
extern char rulesfile[];  // FuzzyCalc.cpp:46 (the rules text)
extern int rules_loaded;  // FuzzyCalc.cpp:44
#include "transforms/wavelet/Symlet5.h"
#include "transforms/wavelet/DataGrid2D.h"

void TNx(double *x, double *out, int N, int height);
void getChCoeff(double *Im, double *out, double *Tj, int N, int width, int height);

extern "C" {

/* The reference's fuzzy colour classifier (colors/FuzzyCalc.cpp:802-831).
 * Its own first-call loader leaves color_functions[] unterminated:
 * LoadColorsFunctions (FuzzyCalc.cpp:582-622) returns on reaching "rules:"
 * before writing the terminating entry, so color_value (625-651) reads past
 * the array for any hue its class does not cover (undefined behaviour: a
 * standalone build segfaults).  The rules are loaded here by the
 * reference's own functions with the rules section hidden while the colour
 * functions load, so that LoadColorsFunctions runs to the end of the text
 * and writes its terminator; then FindColor runs on the loaded tables. */
long ref_find_color(int h, int s, int v) {
  if (!rules_loaded) {
    SetColors();
    char *rules = strstr(rulesfile, "rules:");
    const char saved = *rules;
    *rules = '\0';
    LoadColorsFunctions(rulesfile);
    *rules = saved;
    LoadRules(rules);
    rules_loaded = 1;
  }
  return FindColor((short)h, (short)s, (short)v, nullptr);
}
void ref_find_color_many(const unsigned char *hsv, long n, long *out) {
  for (long i = 0; i < n; i++) out[i] = ref_find_color(hsv[3 * i], hsv[3 * i + 1], hsv[3 * i + 2]);
}

int ref_texture(const unsigned char *grays, int rows, int cols, int angle, double *out14) {
  std::vector<u_int8_t *> rowp(rows);
  std::vector<u_int8_t> buf(grays, grays + (size_t)rows * cols);
  for (int r = 0; r < rows; r++) rowp[r] = &buf[(size_t)r * cols];
  TEXTURE *t = Extract_Texture_Features(1, angle, rowp.data(), rows, cols);
  if (!t) return -1;
  double v[14] = {t->ASM, t->contrast, t->correlation, t->variance, t->IDM, t->sum_avg, t->sum_var,
                  t->sum_entropy, t->entropy, t->diff_var, t->diff_entropy, t->meas_corr1,
                  t->meas_corr2, t->max_corr_coef};
  memcpy(out14, v, sizeof v);
  free(t);
  return 0;
}

int ref_eigen(const double *Q, int n, double *re, double *im) {
  Eigen::Matrix<double, Eigen::Dynamic, Eigen::Dynamic, Eigen::RowMajor> m(n, n);
  for (int i = 0; i < n; i++) for (int j = 0; j < n; j++) m(i, j) = Q[(size_t)i * n + j];
  Eigen::VectorXcd ev = m.eigenvalues();
  for (int i = 0; i < n; i++) { re[i] = ev(i).real(); im[i] = ev(i).imag(); }
  return 0;
}

int ref_radon(const double *pix_colmajor, int M, int N, int rFirst, int rSize, double *out4xr) {
  double theta[4] = {0, 45, 90, 135};
  std::vector<double> in(pix_colmajor, pix_colmajor + (size_t)M * N);
  memset(out4xr, 0, sizeof(double) * 4 * rSize);
  radon(out4xr, in.data(), theta, M, N, (N - 1) / 2, (M - 1) / 2, 4, rFirst, rSize);
  return 0;
}

int ref_wavelet(const double *in, int h, int w, double *out, int *oh, int *ow) {
  DataGrid *grid = new DataGrid2D(w, h, -1);
  for (int y = 0; y < h; y++) for (int x = 0; x < w; x++) grid->setData(x, y, -1, in[(size_t)y * w + x]);
  Symlet5 *s = new Symlet5(0, 1);
  s->transform2D(grid);
  *ow = grid->getX(); *oh = grid->getY();
  for (int y = 0; y < *oh; y++) for (int x = 0; x < *ow; x++) out[(size_t)y * (*ow) + x] = grid->getData(x, y, -1);
  delete s; delete grid;
  return 0;
}

/* first contraction of Chebyshev2D: out[h][N] from rows of in[h][w] */
int ref_cheb1d(const double *in, int h, int w, int N, double *out) {
  std::vector<double> x(w), T((size_t)w * N), im(in, in + (size_t)h * w);
  for (int a = 0; a < w; a++) x[a] = 2 * (double)(a + 1) / (double)w - 1;
  TNx(x.data(), T.data(), N, w);
  getChCoeff(im.data(), out, T.data(), N, w, h);
  return 0;
}
}
