// wndcharm.hpp -- C++ host API of the MI355X feature engine, with the class
// names, member names and argument meaning of the reference's operator API
// for the feature-extraction path:
//
//   ImageMatrix                     src/cmatrix.h:160-260 (FP64 pixel plane; width, height,
//                                   data_ptr / writable_data_ptr)
//   ImageTransform::execute         src/ImageTransforms.h:45-51
//   FeatureAlgorithm::execute       src/FeatureAlgorithms.h:41-51 (n_features, execute)
//   FeatureComputationPlan          src/Tasks.h:272-313 (add, finalize, isFinalized, n_features,
//                                   feature_vec_type, getFGoffset, getFeatureNameByIndex)
//   StdFeatureComputationPlans      src/Tasks.h:356-372 (getFeatureSet, getFeatureSetLong,
//                                   getFeatureSetColor, getFeatureSetLongColor)
//   FeatureComputationPlanExecutor  src/Tasks.h:314-333, run() src/Tasks.cpp:275-295
//   ComputationTaskInstances lookup src/FeatureNames.cpp:251-296 (tasks by registered name)
//
// Everything runs on the GPU through the C ABI of include/wndgpu.h; nothing
// here computes features on the host.  Errors throw wndgpu::Error (the
// reference asserts or silently substitutes an EmptyTransform for unknown
// names; here an unknown name is an error).
//
// Header-only; link with -lwndgpu (wnd-charm_amd/lib).
#pragma once
#include <wndgpu.h>

#include <algorithm>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace wndgpu {

class Error : public std::runtime_error {
 public:
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// The process-wide engine context on device WNDGPU_DEVICE (default 0).
inline wndgpu_ctx *context() {
  static std::once_flag once;
  static wndgpu_ctx *ctx = nullptr;
  static int rc = 0;
  std::call_once(once, [] {
    const char *d = std::getenv("WNDGPU_DEVICE");
    rc = wndgpu_create(d ? std::atoi(d) : 0, &ctx);
  });
  if (rc != WNDGPU_OK) throw Error(rc, "wndgpu_create failed: no usable MI355X GPU");
  return ctx;
}

inline void check(int rc, const std::string &what) {
  if (rc != WNDGPU_OK) throw Error(rc, what + ": " + wndgpu_last_error(context()));
}

// One FP64 row-major pixel plane (the grayscale plane of the reference's
// ImageMatrix; colour planes are not part of this engine).
class ImageMatrix {
 public:
  unsigned int width = 0, height = 0;
  ImageMatrix() = default;
  ImageMatrix(unsigned int w, unsigned int h) { allocate(w, h); }
  void allocate(unsigned int w, unsigned int h) {
    width = w;
    height = h;
    pix_.assign((size_t)w * h, 0.0);
  }
  const double *data_ptr() const { return pix_.data(); }
  double *writable_data_ptr() { return pix_.data(); }
  double &operator()(unsigned int y, unsigned int x) { return pix_[(size_t)y * width + x]; }
  double operator()(unsigned int y, unsigned int x) const { return pix_[(size_t)y * width + x]; }

 private:
  std::vector<double> pix_;
};

class ComputationTask {
 public:
  std::string name;
  virtual ~ComputationTask() = default;

 protected:
  explicit ComputationTask(const std::string &s) : name(s) {}
};

// Transform by registered name: "Fourier", "Chebyshev", "Wavelet", "Edge".
class ImageTransform : public ComputationTask {
 public:
  explicit ImageTransform(const std::string &s) : ComputationTask(s) {
    if (s != "Fourier" && s != "Chebyshev" && s != "Wavelet" && s != "Edge")
      throw Error(WNDGPU_E_ARG, "unknown image transform '" + s + "'");
  }
  virtual void execute(const ImageMatrix &matrix_IN, ImageMatrix &matrix_OUT) const {
    const int h = (int)matrix_IN.height, w = (int)matrix_IN.width;
    // largest output of the four transforms (Wavelet pads by the filter length)
    const size_t cap = (size_t)(2 * ((h + 9) / 2)) * (size_t)(2 * ((w + 9) / 2));
    std::vector<double> buf(cap > (size_t)h * w ? cap : (size_t)h * w);
    int oh = 0, ow = 0;
    check(wndgpu_transform(context(), name.c_str(), matrix_IN.data_ptr(), h, w, buf.data(), &oh, &ow),
          "ImageTransform '" + name + "'");
    matrix_OUT.allocate((unsigned)ow, (unsigned)oh);
    std::copy(buf.begin(), buf.begin() + (size_t)oh * ow, matrix_OUT.writable_data_ptr());
  }
};

// Feature algorithm by registered name ("Zernike Coefficients", ...).
class FeatureAlgorithm : public ComputationTask {
 public:
  int n_features;
  explicit FeatureAlgorithm(const std::string &s) : ComputationTask(s), n_features(wndgpu_algorithm_n_features(s.c_str())) {
    if (n_features < 0) throw Error(WNDGPU_E_ARG, "unknown feature algorithm '" + s + "'");
  }
  virtual std::vector<double> execute(const ImageMatrix &IN_matrix) const {
    std::vector<double> out((size_t)n_features);
    int n = 0;
    check(wndgpu_algorithm(context(), name.c_str(), IN_matrix.data_ptr(), (int)IN_matrix.height,
                           (int)IN_matrix.width, out.data(), &n),
          "FeatureAlgorithm '" + name + "'");
    out.resize((size_t)n);
    return out;
  }
};

// Registry lookups by name (one instance per name, created on first use).
inline const ImageTransform *getImageTransformByName(const std::string &name) {
  static std::mutex mu;
  static std::map<std::string, std::unique_ptr<ImageTransform>> reg;
  std::lock_guard<std::mutex> lk(mu);
  auto &p = reg[name];
  if (!p) p.reset(new ImageTransform(name));
  return p.get();
}
inline const FeatureAlgorithm *getFeatureAlgorithmByName(const std::string &name) {
  static std::mutex mu;
  static std::map<std::string, std::unique_ptr<FeatureAlgorithm>> reg;
  std::lock_guard<std::mutex> lk(mu);
  auto &p = reg[name];
  if (!p) p.reset(new FeatureAlgorithm(name));
  return p.get();
}

// A plan: feature groups ("Haralick Textures (Fourier ())") added by name,
// then finalize(), which builds the shared-prefix transform DAG and assigns
// each group its column offset.
class FeatureComputationPlan {
 public:
  std::string name;
  size_t n_features = 0;
  int feature_vec_type = 0;

  explicit FeatureComputationPlan(const std::string &name_in) : name(name_in) {}
  FeatureComputationPlan(const FeatureComputationPlan &) = delete;
  FeatureComputationPlan &operator=(const FeatureComputationPlan &) = delete;
  virtual ~FeatureComputationPlan() {
    if (plan_) wndgpu_plan_destroy(plan_);
  }

  virtual void add(const std::string &FGname) {
    if (plan_) throw Error(WNDGPU_E_ARG, "FeatureComputationPlan '" + name + "' is finalized");
    groups_.push_back(FGname);
  }
  void finalize() {
    if (plan_) return;
    std::vector<const char *> g;
    for (const auto &s : groups_) g.push_back(s.c_str());
    wndgpu_plan *p = nullptr;
    const int rc = wndgpu_plan_from_groups(g.data(), (int)g.size(), &p);
    if (rc != WNDGPU_OK) throw Error(rc, "FeatureComputationPlan '" + name + "': bad feature group name");
    adopt(p);
  }
  bool isFinalized() const { return plan_ != nullptr; }
  // column where a feature group starts, or size_t(-1) (as the reference)
  size_t getFGoffset(const std::string &FGname) const {
    const int o = plan_ ? wndgpu_plan_group_offset(plan_, FGname.c_str()) : -1;
    return o < 0 ? (size_t)-1 : (size_t)o;
  }
  const std::string &getFeatureNameByIndex(size_t offset) const {
    if (!plan_ || offset >= n_features) throw Error(WNDGPU_E_ARG, "feature index out of range");
    return names_[offset];
  }
  const wndgpu_plan *handle() const { return plan_; }

  // wraps a plan built by the engine (standard plans)
  void adopt(wndgpu_plan *p) {
    plan_ = p;
    n_features = (size_t)wndgpu_plan_n_features(p);
    names_.resize(n_features);
    for (size_t i = 0; i < n_features; i++) names_[i] = wndgpu_plan_feature_name(p, (int)i);
  }

 private:
  std::vector<std::string> groups_;
  std::vector<std::string> names_;
  wndgpu_plan *plan_ = nullptr;
};

struct StdFeatureComputationPlans {
  enum feature_vec_types { fv_unknown = 0, fv_short = 1, fv_long = 2, fv_short_color = 3, fv_long_color = 4 };
  static const FeatureComputationPlan *getFeatureSet() { return get(0, fv_short, "Standard Feature Set"); }
  static const FeatureComputationPlan *getFeatureSetLong() { return get(1, fv_long, "Long Feature Set"); }
  static const FeatureComputationPlan *getFeatureSetColor() { return get(2, fv_short_color, "Color Feature Set"); }
  static const FeatureComputationPlan *getFeatureSetLongColor() {
    return get(3, fv_long_color, "Long Color Feature Set");
  }

 private:
  static const FeatureComputationPlan *get(int set, int type, const char *nm) {
    static std::mutex mu;
    static std::unique_ptr<FeatureComputationPlan> plans[4];
    std::lock_guard<std::mutex> lk(mu);
    if (!plans[set]) {
      wndgpu_plan *p = nullptr;
      const int rc = wndgpu_plan_standard(set, &p);
      if (rc != WNDGPU_OK) throw Error(rc, "wndgpu_plan_standard failed");
      plans[set].reset(new FeatureComputationPlan(nm));
      plans[set]->adopt(p);
      plans[set]->feature_vec_type = type;
    }
    return plans[set].get();
  }
};

// Runs a finalized plan on one image (run) or on a batch (run_batch, the
// engine's native form: every transform and feature group is one launch
// over all images).  Row r of feature_mat is feature_mat[n_features*r ...].
class FeatureComputationPlanExecutor {
 public:
  const FeatureComputationPlan *plan;
  explicit FeatureComputationPlanExecutor(const FeatureComputationPlan *plan_in) : plan(plan_in) {
    if (!plan || !plan->isFinalized()) throw Error(WNDGPU_E_ARG, "executor needs a finalized plan");
  }
  virtual void run(const ImageMatrix *source_mat, std::vector<double> &feature_mat_in, size_t dest_row) {
    run_batch(std::vector<const ImageMatrix *>{source_mat}, feature_mat_in, dest_row);
  }
  void run_batch(const std::vector<const ImageMatrix *> &sources, std::vector<double> &feature_mat_in,
                 size_t first_row) {
    const size_t nf = plan->n_features;
    if (feature_mat_in.size() < nf * (first_row + sources.size()))
      throw Error(WNDGPU_E_ARG, "feature matrix too small for the destination rows");
    std::vector<wndgpu_image> imgs(sources.size());
    for (size_t i = 0; i < sources.size(); i++)
      imgs[i] = wndgpu_image{sources[i]->data_ptr(), (int)sources[i]->height, (int)sources[i]->width,
                             (int)sources[i]->width, WNDGPU_F64};
    std::vector<int> status(sources.size());
    check(wndgpu_compute(context(), plan->handle(), imgs.data(), (int)imgs.size(),
                         feature_mat_in.data() + nf * first_row, status.data()),
          "FeatureComputationPlanExecutor::run ('" + plan->name + "')");
  }
};

}  // namespace wndgpu
