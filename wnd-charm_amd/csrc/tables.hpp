// Host-computed geometry tables (see tables.cpp).
#pragma once
#include <vector>

namespace wnd {

std::vector<double> cheb_basis(int len, int N);
void fft_tables(int n, std::vector<double> &tw, std::vector<int> &radix);

struct CfTable {
  int NN = 0;
  std::vector<int> kk;
  std::vector<double> tn, fc, fs;
};
CfTable cf_table(int w, int h);

std::vector<double> ds_positions(unsigned len, double dx);
std::vector<double> ds_dense(unsigned len, double dx, unsigned nout);
std::vector<double> gabor_kernels();
std::vector<double> zernike_h();
std::vector<double> radon_trig();
std::vector<double> comb_offsets(int m);

}  // namespace wnd
