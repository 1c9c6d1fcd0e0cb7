// Drives the C++ host API the way the reference's signatures::compute_plan
// (src/wndchrm_src/signatures.cpp:167-175) and a transform / algorithm
// caller would:
//   facade_check <image.f64> <height> <width> <out.f64>
// reads a row-major FP64 plane, writes [long-set vector (2919)] [Zernike (72)]
// [Fourier plane (h*w)] [short-set vector (1059)] as FP64 to <out.f64>.
// Exit status 0 on success; any engine error is printed and returns 1.
#include <cstdio>
#include <fstream>
#include <iostream>

#include "wndcharm.hpp"

int main(int argc, char **argv) {
  if (argc != 5) {
    std::fprintf(stderr, "usage: %s image.f64 height width out.f64\n", argv[0]);
    return 2;
  }
  const unsigned h = (unsigned)std::atoi(argv[2]), w = (unsigned)std::atoi(argv[3]);
  wndgpu::ImageMatrix img(w, h);
  std::ifstream in(argv[1], std::ios::binary);
  in.read((char *)img.writable_data_ptr(), (std::streamsize)((size_t)w * h * sizeof(double)));
  if (!in) {
    std::fprintf(stderr, "cannot read %s\n", argv[1]);
    return 2;
  }
  try {
    const wndgpu::FeatureComputationPlan *plan = wndgpu::StdFeatureComputationPlans::getFeatureSetLong();
    std::vector<double> sig(plan->n_features);
    wndgpu::FeatureComputationPlanExecutor exec(plan);
    exec.run(&img, sig, 0);

    std::vector<double> zern = wndgpu::getFeatureAlgorithmByName("Zernike Coefficients")->execute(img);

    wndgpu::ImageMatrix four;
    wndgpu::getImageTransformByName("Fourier")->execute(img, four);

    // a custom plan built group by group, as FeatureComputationPlan::add does
    wndgpu::FeatureComputationPlan custom("short, by groups");
    const wndgpu::FeatureComputationPlan *shortp = wndgpu::StdFeatureComputationPlans::getFeatureSet();
    std::vector<std::string> groups;
    for (size_t i = 0; i < shortp->n_features; i++) {
      const std::string &nm = shortp->getFeatureNameByIndex(i);
      const std::string g = nm.substr(0, nm.rfind(" ["));
      if (groups.empty() || groups.back() != g) groups.push_back(g);
    }
    for (const auto &g : groups) custom.add(g);
    custom.finalize();
    std::vector<double> sig2(custom.n_features);
    wndgpu::FeatureComputationPlanExecutor(&custom).run(&img, sig2, 0);

    std::ofstream out(argv[4], std::ios::binary);
    out.write((const char *)sig.data(), (std::streamsize)(sig.size() * sizeof(double)));
    out.write((const char *)zern.data(), (std::streamsize)(zern.size() * sizeof(double)));
    out.write((const char *)four.data_ptr(), (std::streamsize)((size_t)four.width * four.height * sizeof(double)));
    out.write((const char *)sig2.data(), (std::streamsize)(sig2.size() * sizeof(double)));
    std::printf("facade_check: long %zu, zernike %zu, fourier %ux%u, short-by-groups %zu\n", sig.size(), zern.size(),
                four.height, four.width, sig2.size());
  } catch (const wndgpu::Error &e) {
    std::fprintf(stderr, "wndgpu error %d: %s\n", e.code, e.what());
    return 1;
  }
  return 0;
}
