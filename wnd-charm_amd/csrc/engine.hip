// Host orchestration of the feature DAG on one MI355X and the C ABI
// (include/wndgpu.h).
//
// A batch of images runs node by node through the transform DAG of a plan
// (FeatureComputationPlanExecutor::run, src/Tasks.cpp:185-311): every node
// is one batched transform launch whose output planes stay resident in HBM
// for the whole batch, and every feature group on a node is one batched
// launch writing its columns of the [n_images][n_features] result.  All
// device memory of a batch comes from one arena; all kernel descriptors are
// staged in pinned host memory and uploaded with a single copy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <atomic>
#include <chrono>
#include <climits>
#include <condition_variable>
#include <deque>
#include <future>
#include <mutex>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/wndgpu.h"
#include "kernels.hpp"
#include "tables.hpp"

namespace wnd {
void launch_norm_train(double *data, int n, int f, double *mins, double *maxs, hipStream_t s);
void launch_norm_test(double *vec, int nvec, int f, const double *mins, const double *maxs, hipStream_t s);
void launch_fisher(const double *data, const int *cls, int n, int f, int C, double used, double *w, hipStream_t s);
void launch_py_normalize(double *M, int n, int f, int flags, double *mins, double *maxs, double *means,
                         double *stdevs, int *flag, hipStream_t s);
void launch_py_fisher(const double *data, const int *cls, int n, int f, int C, double *w, hipStream_t s);
void launch_py_wnd5(const double *train, const double *w, const double *test, const int *rows, const int *start,
                    int n, int f, int C, int ntest, double *dist, double *sims, double *cmean, double *prob,
                    double *norm, long *pred, hipStream_t s);
void launch_wnd5(const double *train, const int *cls, const double *w, const double *test, int n, int f, int C,
                 int ntest, double *dist, double *prob, double *norm, long *pred, hipStream_t s);
void launch_wnn(const double *train, const int *cls, const double *w, const double *test, int n, int f, int C,
                int ntest, double *dist, double *prob, double *norm, long *pred, hipStream_t s);
}  // namespace wnd

using namespace wnd;

// ============================================================ names / plans
// "Color Transform" is the name the reference's standard colour plans use
// (Tasks.cpp:448-449); no transform registers it, so the reference runs an
// EmptyTransform, a copy of the image (FeatureNames.cpp:276-285,
// ImageTransforms.cpp:51-55).  T_COLOR_COPY is that identity.
enum { T_FOURIER = 0, T_CHEBYSHEV = 1, T_WAVELET = 2, T_EDGE = 3, T_COLOR = 4, T_HUE = 5, T_COLOR_COPY = 6, T_COUNT = 7 };
enum {
  A_CHEB_FOURIER = 0, A_CHEBYSHEV, A_ZERNIKE, A_HARALICK, A_MSH, A_TAMURA, A_COMB, A_RADON,
  A_FRACTAL, A_PIXSTATS, A_EDGE, A_OTSU, A_INV_OTSU, A_GABOR, A_GINI, A_COLOR_HIST, A_COUNT
};
static const int kNFeat[A_COUNT] = {32, 32, 72, 28, 24, 6, 48, 12, 20, 5, 28, 34, 34, 7, 1, 20};
// registered names (src/FeatureAlgorithms.cpp:54-489, src/ImageTransforms.cpp:57-103)
static const char *kAlgNames[A_COUNT] = {
    "Chebyshev-Fourier Coefficients", "Chebyshev Coefficients", "Zernike Coefficients", "Haralick Textures",
    "Multiscale Histograms", "Tamura Textures", "Comb Moments", "Radon Coefficients", "Fractal Features",
    "Pixel Intensity Statistics", "Edge Features", "Otsu Object Features", "Inverse-Otsu Object Features",
    "Gabor Textures", "Gini Coefficient", "Color Histogram"};
static const char *kTrNames[T_COUNT] = {"Fourier", "Chebyshev", "Wavelet", "Edge", "Color", "Hue", "Color Transform"};

static std::string trim(const std::string &s) {
  size_t a = s.find_first_not_of(" \t\f\v\n\r"), b = s.find_last_not_of(" \t\f\v\n\r");
  return a == std::string::npos ? "" : s.substr(a, b - a + 1);
}
static int alg_index(const std::string &n) {
  for (int i = 0; i < A_COUNT; i++) if (n == kAlgNames[i]) return i;
  return -1;
}
static int tr_index(const std::string &n) {
  for (int i = 0; i < T_COUNT; i++) if (n == kTrNames[i]) return i;
  return -1;
}
static std::string node_text(const std::vector<int> &chain) {
  std::string s = "()";
  for (int t : chain) s = "(" + std::string(kTrNames[t]) + " " + s + ")";
  return s;
}

struct Group {
  std::vector<int> chain;  // innermost first
  int alg;
  int offset;
  std::string name;
};

static std::atomic<long> g_plan_uid{0};
struct wndgpu_plan {
  long uid = ++g_plan_uid;  // identity of the plan in batch-graph signatures
  std::vector<Group> groups;
  int n_features = 0;
  bool color = false;  // some group reads the colour plane (RGB images only)
  // arithmetic of the basis contractions of Chebyshev Coefficients and the
  // Gabor filter bank: 0 FP64 (the plain form), 1 FP32, 2 BF16x3 on MFMA
  int precision = 0;
  std::vector<std::string> names;
  std::vector<std::vector<int>> nodes;  // DAG nodes, parents before children
};

// "Alg (T2 (T1 ()))" -> alg, chain {T1, T2} (FeatureNames::getGroupByName, FeatureNames.cpp:309-396)
static bool parse_group(const std::string &name, int *alg, std::vector<int> *chain) {
  size_t lp = name.find('(');
  if (lp == std::string::npos) return false;
  *alg = alg_index(trim(name.substr(0, lp)));
  if (*alg < 0) return false;
  std::string rest = trim(name.substr(lp));
  std::vector<int> outer_first;
  while (true) {
    if (rest.size() < 2 || rest.front() != '(' || rest.back() != ')') return false;
    std::string inner = trim(rest.substr(1, rest.size() - 2));
    if (inner.empty()) break;
    size_t p = inner.find('(');
    if (p == std::string::npos) return false;
    const int t = tr_index(trim(inner.substr(0, p)));
    if (t < 0) return false;
    outer_first.push_back(t);
    rest = trim(inner.substr(p));
  }
  chain->assign(outer_first.rbegin(), outer_first.rend());
  // the colour plane belongs to the loaded image: Color and Hue apply to it
  // directly, and the Color Histogram reads it
  for (size_t k = 0; k < chain->size(); k++)
    if (((*chain)[k] == T_COLOR || (*chain)[k] == T_HUE) && k != 0) return false;
  if (*alg == A_COLOR_HIST && !chain->empty()) return false;
  return true;
}

static void plan_add(wndgpu_plan *p, int alg, const std::vector<int> &chain) {
  Group g{chain, alg, p->n_features, std::string(kAlgNames[alg]) + " " + node_text(chain)};
  for (int i = 0; i < kNFeat[alg]; i++) p->names.push_back(g.name + " [" + std::to_string(i) + "]");
  p->n_features += kNFeat[alg];
  p->groups.push_back(g);
  if (alg == A_COLOR_HIST || (!chain.empty() && (chain[0] == T_COLOR || chain[0] == T_HUE))) p->color = true;
  std::vector<int> key;
  auto has = [&](const std::vector<int> &k) { return std::find(p->nodes.begin(), p->nodes.end(), k) != p->nodes.end(); };
  if (!has(key)) p->nodes.push_back(key);
  for (int t : chain) {
    key.push_back(t);
    if (!has(key)) p->nodes.push_back(key);
  }
}

static void plan_finalize(wndgpu_plan *p) {
  std::stable_sort(p->nodes.begin(), p->nodes.end(),
                   [](const std::vector<int> &a, const std::vector<int> &b) { return a.size() < b.size(); });
}

// StdFeatureComputationPlans (src/Tasks.cpp:313-508): set 0 = short, 1 =
// long, 2 = short colour, 3 = long colour
static wndgpu_plan *make_standard(int set) {
  wndgpu_plan *p = new wndgpu_plan;
  const int F = T_FOURIER, C = T_CHEBYSHEV, W = T_WAVELET, E = T_EDGE, H = T_HUE, CT = T_COLOR_COPY;
  auto A = [&](std::vector<int> ch) { for (int a : {A_EDGE, A_OTSU, A_INV_OTSU, A_GABOR}) plan_add(p, a, ch); };
  auto B = [&](std::vector<int> ch) { for (int a : {A_CHEB_FOURIER, A_CHEBYSHEV, A_ZERNIKE}) plan_add(p, a, ch); };
  auto Cg = [&](std::vector<int> ch) {
    for (int a : {A_COMB, A_HARALICK, A_MSH, A_TAMURA, A_RADON, A_FRACTAL, A_PIXSTATS, A_GINI}) plan_add(p, a, ch);
  };
  auto colour = [&]() {  // addColorFeatures, Tasks.cpp:445-459
    plan_add(p, A_COLOR_HIST, {});
    B({CT}); Cg({CT});
    B({H}); Cg({H});
    B({H, F}); Cg({H, F});
    B({H, C}); Cg({H, C});
  };
  if (set == 1 || set == 3) {
    A({}); B({}); Cg({});
    if (set == 3) colour();
    B({F}); Cg({F});
    B({W}); Cg({W});
    B({C}); Cg({C});
    Cg({F, C});
    Cg({F, W});
    B({W, F}); Cg({W, F});
    Cg({C, F});
    Cg({W, C});
    B({E}); Cg({E});
    B({E, F}); Cg({E, F});
    B({E, W}); Cg({E, W});
  } else {
    plan_add(p, A_CHEB_FOURIER, {}); plan_add(p, A_CHEB_FOURIER, {F});
    plan_add(p, A_CHEBYSHEV, {}); plan_add(p, A_CHEBYSHEV, {F});
    plan_add(p, A_COMB, {}); plan_add(p, A_COMB, {C}); plan_add(p, A_COMB, {F, C}); plan_add(p, A_COMB, {F});
    plan_add(p, A_COMB, {W}); plan_add(p, A_COMB, {F, W});
    plan_add(p, A_EDGE, {}); plan_add(p, A_OTSU, {}); plan_add(p, A_INV_OTSU, {}); plan_add(p, A_GABOR, {});
    plan_add(p, A_HARALICK, {}); plan_add(p, A_HARALICK, {C}); plan_add(p, A_HARALICK, {F, C});
    plan_add(p, A_HARALICK, {F}); plan_add(p, A_HARALICK, {W}); plan_add(p, A_HARALICK, {F, W});
    plan_add(p, A_MSH, {}); plan_add(p, A_MSH, {C}); plan_add(p, A_MSH, {F, C}); plan_add(p, A_MSH, {F});
    plan_add(p, A_MSH, {W}); plan_add(p, A_MSH, {F, W});
    plan_add(p, A_RADON, {}); plan_add(p, A_RADON, {C}); plan_add(p, A_RADON, {F, C}); plan_add(p, A_RADON, {F});
    plan_add(p, A_TAMURA, {}); plan_add(p, A_TAMURA, {C}); plan_add(p, A_TAMURA, {F, C});
    plan_add(p, A_TAMURA, {F}); plan_add(p, A_TAMURA, {W}); plan_add(p, A_TAMURA, {F, W});
    plan_add(p, A_ZERNIKE, {}); plan_add(p, A_ZERNIKE, {F});
    if (set == 2) colour();
  }
  plan_finalize(p);
  return p;
}

// ============================================================ the context
struct DevTable {
  void *p = nullptr;
  size_t bytes = 0;
  long used = 0;  // last batch that read it
};

// A fixed pool of host threads for size-dependent tables computed ahead of
// the batch that needs them (wndgpu_prefetch_tables).  Tasks only read their
// own arguments and fulfil a promise: they never touch the context.
struct TablePool {
  std::mutex mu;
  std::condition_variable cv;
  std::deque<std::function<void()>> q;
  std::vector<std::thread> th;
  bool stop = false;
  explicit TablePool(unsigned n) {
    for (unsigned t = 0; t < n; t++)
      th.emplace_back([this]() {
        for (;;) {
          std::function<void()> f;
          {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [this]() { return stop || !q.empty(); });
            if (stop && q.empty()) return;
            f = std::move(q.front());
            q.pop_front();
          }
          f();
        }
      });
  }
  void push(std::function<void()> f) {
    {
      std::lock_guard<std::mutex> lk(mu);
      q.push_back(std::move(f));
    }
    cv.notify_one();
  }
  ~TablePool() {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    cv.notify_all();
    for (auto &t : th) t.join();
  }
};

struct CfDev {
  DevTable kk, tn, fc, fs;
  int np = 0, NN = 0;
  long used = 0;
  size_t bytes() const { return kk.bytes + tn.bytes + fc.bytes + fs.bytes; }
};

struct NodeRun {
  std::vector<int> chain;
  std::vector<int> h, w;
  Img *imgs = nullptr;    // device descriptor array
  Stats *stats = nullptr; // device stats array
  std::vector<double *> planes;
  std::vector<unsigned char *> hsv;            // colour planes (root of RGB images)
  unsigned char *const *hsvd = nullptr;        // device copy of hsv
  double ms = 0;
};

struct wndgpu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // The Haralick chain (co-occurrence .. Schur) runs on `side`, forked from
  // `stream` once every transform node is issued and joined back before the
  // batch completes: its memory- and latency-bound eigen kernels overlap
  // the arithmetic-bound feature kernels, which alternate between `stream`
  // and `fstream` so that the tail of one launch overlaps the next.  All
  // streams have the default priority.
  hipStream_t side = nullptr;
  hipStream_t fstream = nullptr;  // every other non-Haralick feature launch
  hipEvent_t fork_ev = nullptr, join_ev = nullptr, fjoin_ev = nullptr;
  void *scratch = nullptr;  // this device's copy of `scratch` (wndgpu_ad_wnd_forward_noalloc)
  bool serial = false;  // WNDGPU_SERIAL: every launch on `stream` (per-kernel timing without overlap)
  std::string err;
  // arena
  char *dev = nullptr;
  size_t dev_cap = 0, dev_used = 0;
  char *host = nullptr;   // pinned descriptor staging
  char *hdesc = nullptr;  // device copy of the staging
  size_t host_cap = 0, host_used = 0;
  // caches (device)
  std::map<std::pair<int, int>, DevTable> cheb;  // (len, N)
  std::map<int, std::pair<DevTable, std::vector<int>>> fft;
  std::map<std::pair<int, int>, CfDev> cf;
  std::map<std::pair<int, double>, DevTable> dspos;
  std::map<std::tuple<int, double, int>, DevTable> dsdense;  // (len, dx, nout): dense Downsample passes
  std::map<int, DevTable> comb;
  DevTable gabor, zern, rtrig;
  DevTable color_lut;  // fuzzy colour class of every (h, s, v) byte triple, built on first use
  // Size-dependent bases (Chebyshev, Chebyshev-Fourier geometry) take libm
  // calls on the host per new image size: the sizing pass of a batch lists
  // the ones it needs, the missing ones are computed on host threads
  // together, and least recently used ones beyond table_budget bytes are
  // freed between batches.
  std::set<std::pair<int, int>> need_cheb, need_cf;
  long batch_id = 0;
  size_t table_budget = (size_t)16 << 30;
  // tables queued by wndgpu_prefetch_tables: taken by the batch that needs
  // them, dropped two batches after they were queued otherwise
  std::unique_ptr<TablePool> pool;
  template <class T>
  struct Pending {
    std::future<T> f;
    long queued;
  };
  std::map<std::pair<int, int>, Pending<DevTable>> pend_cheb;
  std::map<std::pair<int, int>, Pending<CfDev>> pend_cf;
  hipStream_t ustream = nullptr;  // non-blocking stream of the pool's table uploads
  // wall times of the last compute call (wndgpu_last_call_times)
  double t_host_ms = 0, t_table_ms = 0, t_device_ms = 0;
  // Gabor kernels for the reduced-precision forms: FP32, bf16 hi / mid / lo
  // planes, and per filter the sum of |re| + |im| of the taps (error bounds)
  DevTable gabor_f32, gabor_bhi, gabor_bmd, gabor_blo, gabor_wsum;
  // Batch graphs (wndgpu_set_graph_mode): the launch sequence of a batch
  // whose signature (plan, geometry, arena and table state) was seen before
  // is captured into a hipGraph and replayed; mode 2 also times the two
  // launch schedules (side stream / one stream) and keeps the faster.
  struct BatchGraph {
    hipGraphExec_t exec[2] = {nullptr, nullptr};  // [0] side-stream schedule, [1] one stream
    float ms[2] = {-1.f, -1.f};
    int pick = -1;      // schedule in use once both are timed (mode 2)
    int pending = -1;   // schedule whose events of the previous call are still to be read
    long used = 0;
  };
  int graph_mode = 0;
  std::map<std::string, BatchGraph> graphs;
  std::set<std::string> graph_seen, graph_failed;
  long graph_clock = 0, graph_captures = 0, graph_replays = 0;
  hipEvent_t gev0 = nullptr, gev1 = nullptr;
  std::vector<std::pair<std::string, double>> times;
  // per-launch timing (WNDGPU_TIMING): an event pair around every launch,
  // read after the compute call has synchronised
  std::vector<hipEvent_t> evpool;
  std::vector<std::string> tnames;
  // Haralick entries of the last compute: tone counts and Schur-iteration flops
  int *har_ng = nullptr;
  double *har_work = nullptr;
  int har_n = 0;
  // Chebyshev-Fourier entries of the last compute: open[k] = 1 where the
  // matrix-core histogram was left uncertified and recomputed in the
  // reference's order.  cf_slack (WNDGPU_CF_SLACK at creation, >= 1)
  // scales the certification bound.
  int *cf_open = nullptr;
  int cf_n = 0;
  double cf_slack = 1;
  // Gabor energies of the plain form: FP64 fused multiply-adds whose counts
  // are certified (the default), or the reference's separate multiply and
  // add throughout (gabor_direct, WNDGPU_GABOR_DIRECT=1 at creation).
  // gabor_slack (WNDGPU_GABOR_SLACK, >= 1) scales the certification bound.
  bool gabor_direct = false;
  double gabor_slack = 1;
  // Radon projections: order-free sums whose histogram counts are
  // certified (the default), or the reference's sequential sum per bin
  // (radon_direct, WNDGPU_RADON_DIRECT=1 at creation).  radon_slack
  // (WNDGPU_RADON_SLACK, >= 1) scales the certification bound.
  bool radon_direct = false;
  double radon_slack = 1;
  bool sized = false;
};

static int fail(wndgpu_ctx *c, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}
#define HIPCHK(ctx, x)                                                                              \
  do {                                                                                              \
    hipError_t e_ = (x);                                                                            \
    if (e_ != hipSuccess) return fail(ctx, WNDGPU_E_DEVICE, "%s: %s", #x, hipGetErrorString(e_));   \
  } while (0)

// Release calls on paths that cannot return an error (teardown, destructors):
// a failure is reported on stderr rather than dropped.
static void release_check(hipError_t e, const char *what) {
  if (e != hipSuccess) fprintf(stderr, "wndgpu: %s failed: %s\n", what, hipGetErrorString(e));
}

// Tables come from the device's stream-ordered pool (hipMallocAsync), so
// evicting one is a hipFreeAsync into the pool, not a synchronising hipFree.
// Uploads run while `stream` is idle (before a batch's launches are issued).
static int upload(wndgpu_ctx *c, DevTable &t, const void *src, size_t bytes) {
  HIPCHK(c, hipMallocAsync(&t.p, bytes ? bytes : 8, c->stream));
  t.bytes = bytes;
  if (bytes) HIPCHK(c, hipMemcpyAsync(t.p, src, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

// --------------------------------------------------------------- arenas
// Two passes over the batch: a sizing pass with null bases, then the real
// pass.  dalloc / halloc are bump allocators over the device arena and the
// pinned staging buffer; hptr_dev maps a staging address to its device copy.
struct Arena {
  wndgpu_ctx *c;
  bool sizing;
  size_t d = 0, h = 0;
  void *dalloc(size_t bytes) {
    size_t off = (d + 255) & ~(size_t)255;
    d = off + bytes;
    return sizing ? (void *)(uintptr_t)(off + 256) : (void *)(c->dev + off);
  }
  template <typename T>
  T *dalloc_n(size_t n) { return (T *)dalloc(n * sizeof(T)); }
  // stage n descriptors; returns the device address the kernels use
  template <typename T>
  T *stage(const std::vector<T> &v) {
    size_t off = (h + 63) & ~(size_t)63;
    h = off + v.size() * sizeof(T);
    if (sizing) return (T *)(uintptr_t)(off + 64);
    memcpy(c->host + off, v.data(), v.size() * sizeof(T));
    return (T *)(c->hdesc + off);
  }
};

// ============================================================ batch execution
// One input image under sampling: OpenImage's crop, Downsample and
// normalisation (cmatrix.cpp:252-294), before its samples are cut.
struct PreImage {
  LoadDesc full;      // the whole image (the channel range of wide RGB samples)
  LoadDesc roi;       // the crop
  int ch, cw;         // crop size
  int dh, dw;         // size after Downsample
  bool ds;            // Downsample applies
  double dx;          // its step, 1 / ratio
  bool norm;          // normalise to mean / stdev
  double mean, stdev, max_range;
};
// One sample: a rotation and tile of a pre-processed image
struct SampleRef {
  int img, rot, x0, y0, h, w;
};

struct Batch {
  int n;  // root entries: images, or samples when sampled
  std::vector<int> h0, w0;
  std::vector<LoadDesc> load;
  bool sampled = false;
  std::vector<PreImage> pre;
  std::vector<SampleRef> smp;
  double *out;  // device [n][nf]
  int nf;
};

static int get_cheb(wndgpu_ctx *c, int len, int N, const double **p) {
  auto key = std::make_pair(len, N);
  auto it = c->cheb.find(key);
  if (it == c->cheb.end()) {
    std::vector<double> b = cheb_basis(len, N);
    DevTable t;
    int r = upload(c, t, b.data(), b.size() * sizeof(double));
    if (r) return r;
    it = c->cheb.emplace(key, t).first;
  }
  it->second.used = c->batch_id;
  *p = (const double *)it->second.p;
  return 0;
}

static int get_fft(wndgpu_ctx *c, int n, FftPlan *pl) {
  auto it = c->fft.find(n);
  if (it == c->fft.end()) {
    std::vector<double> tw;
    std::vector<int> radix;
    fft_tables(n, tw, radix);
    DevTable t;
    int r = upload(c, t, tw.data(), tw.size() * sizeof(double));
    if (r) return r;
    it = c->fft.emplace(n, std::make_pair(t, radix)).first;
  }
  pl->tw = (const double2 *)it->second.first.p;
  pl->n = n;
  pl->nst = (int)it->second.second.size();
  if (pl->nst > 20) return fail(c, WNDGPU_E_ARG, "FFT length %d has too many factors", n);
  for (int i = 0; i < pl->nst; i++) pl->radix[i] = it->second.second[i];
  return 0;
}

static int get_cf(wndgpu_ctx *c, int w, int h, CfDev **out) {
  auto key = std::make_pair(w, h);
  auto it = c->cf.find(key);
  if (it == c->cf.end()) {
    CfTable t = cf_table(w, h);
    CfDev d;
    d.np = (int)t.kk.size();
    d.NN = t.NN;
    int r;
    if ((r = upload(c, d.kk, t.kk.data(), t.kk.size() * sizeof(int)))) return r;
    if ((r = upload(c, d.tn, t.tn.data(), t.tn.size() * sizeof(double)))) return r;
    if ((r = upload(c, d.fc, t.fc.data(), t.fc.size() * sizeof(double)))) return r;
    if ((r = upload(c, d.fs, t.fs.data(), t.fs.size() * sizeof(double)))) return r;
    it = c->cf.emplace(key, d).first;
  }
  it->second.used = c->batch_id;
  *out = &it->second;
  return 0;
}

// Cached batch graphs hold device addresses of the arena and the tables in
// their kernel arguments: they are dropped whenever either is freed.
static void drop_graphs(wndgpu_ctx *c) {
  for (auto &kv : c->graphs)
    for (hipGraphExec_t &e : kv.second.exec)
      if (e) release_check(hipGraphExecDestroy(e), "hipGraphExecDestroy");
  c->graphs.clear();
  c->graph_seen.clear();
}

// The size-dependent tables a batch's sizing pass listed: missing ones are
// computed on up to 16 host threads (glibc libm, as the reference), then
// uploaded; afterwards tables no batch has read recently are freed while the
// cache holds more than table_budget bytes.
// Prefetched tables queued before batch `before` (all for LONG_MAX): their
// tasks are awaited and their device copies freed.
static void drop_pending(wndgpu_ctx *c, long before) {
  auto fr = [c](DevTable &t) {
    if (t.p) release_check(hipFreeAsync(t.p, c->stream), "hipFreeAsync");
    t.p = nullptr;
  };
  for (auto it = c->pend_cheb.begin(); it != c->pend_cheb.end();) {
    if (it->second.queued >= before) { ++it; continue; }
    try { DevTable t = it->second.f.get(); fr(t); } catch (const std::exception &) {}
    it = c->pend_cheb.erase(it);
  }
  for (auto it = c->pend_cf.begin(); it != c->pend_cf.end();) {
    if (it->second.queued >= before) { ++it; continue; }
    try {
      CfDev d = it->second.f.get();
      fr(d.kk); fr(d.tn); fr(d.fc); fr(d.fs);
    } catch (const std::exception &) {}
    it = c->pend_cf.erase(it);
  }
}

// A host table copied to the device on the pool's upload stream (runs on a
// pool thread: no context access; errors are thrown to the awaiting call).
static DevTable upload_on(int device, hipStream_t st, const void *src, size_t bytes) {
  auto chk = [](hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
  };
  DevTable t;
  chk(hipSetDevice(device), "hipSetDevice");
  chk(hipMallocAsync(&t.p, bytes ? bytes : 8, st), "hipMallocAsync");
  t.bytes = bytes;
  hipError_t e = bytes ? hipMemcpyAsync(t.p, src, bytes, hipMemcpyHostToDevice, st) : hipSuccess;
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    release_check(hipFreeAsync(t.p, st), "hipFreeAsync");
    chk(e, "hipMemcpyAsync");
  }
  return t;
}

static CfDev upload_cf_on(int device, hipStream_t st, const CfTable &t) {
  CfDev d;
  d.np = (int)t.kk.size();
  d.NN = t.NN;
  try {
    d.kk = upload_on(device, st, t.kk.data(), t.kk.size() * sizeof(int));
    d.tn = upload_on(device, st, t.tn.data(), t.tn.size() * sizeof(double));
    d.fc = upload_on(device, st, t.fc.data(), t.fc.size() * sizeof(double));
    d.fs = upload_on(device, st, t.fs.data(), t.fs.size() * sizeof(double));
  } catch (...) {
    for (DevTable *x : {&d.kk, &d.tn, &d.fc, &d.fs})
      if (x->p) release_check(hipFreeAsync(x->p, st), "hipFreeAsync");
    throw;
  }
  return d;
}

static double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

static int prepare_tables_impl(wndgpu_ctx *c);
static int prepare_tables(wndgpu_ctx *c) {
  const auto t0 = std::chrono::steady_clock::now();
  const int r = prepare_tables_impl(c);
  c->t_table_ms += ms_since(t0);
  return r;
}

static int prepare_tables_impl(wndgpu_ctx *c) {
  c->batch_id++;
  std::vector<std::pair<int, int>> cheb_keys, cf_keys;
  for (const auto &k : c->need_cheb)
    if (!c->cheb.count(k)) cheb_keys.push_back(k);
  for (const auto &k : c->need_cf)
    if (!c->cf.count(k)) cf_keys.push_back(k);
  const size_t nwork = cheb_keys.size() + cf_keys.size();
  if (nwork) {
    std::vector<std::vector<double>> cheb_v(cheb_keys.size());
    std::vector<CfTable> cf_v(cf_keys.size());
    std::vector<DevTable> cheb_d(cheb_keys.size());  // prefetched: already on the device
    std::vector<CfDev> cf_d(cf_keys.size());
    // keys already queued by a prefetch are awaited, the rest computed here
    std::vector<char> queued(nwork, 0);
    for (size_t j = 0; j < cheb_keys.size(); j++) queued[j] = c->pend_cheb.count(cheb_keys[j]) != 0;
    for (size_t j = 0; j < cf_keys.size(); j++) queued[cheb_keys.size() + j] = c->pend_cf.count(cf_keys[j]) != 0;
    std::atomic<size_t> next{0};
    auto worker = [&]() {
      for (size_t j = next++; j < nwork; j = next++) {
        if (queued[j]) continue;
        if (j < cheb_keys.size()) cheb_v[j] = cheb_basis(cheb_keys[j].first, cheb_keys[j].second);
        else cf_v[j - cheb_keys.size()] = cf_table(cf_keys[j - cheb_keys.size()].first, cf_keys[j - cheb_keys.size()].second);
      }
    };
    const size_t nt = std::min<size_t>(nwork, std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency())));
    std::vector<std::thread> th;
    for (size_t t = 1; t < nt; t++) th.emplace_back(worker);
    worker();
    for (auto &t : th) t.join();
    std::string perr;
    for (size_t j = 0; j < cheb_keys.size(); j++)
      if (queued[j]) {
        try { cheb_d[j] = c->pend_cheb[cheb_keys[j]].f.get(); } catch (const std::exception &e) { perr = e.what(); }
        c->pend_cheb.erase(cheb_keys[j]);
      }
    for (size_t j = 0; j < cf_keys.size(); j++)
      if (queued[cheb_keys.size() + j]) {
        try { cf_d[j] = c->pend_cf[cf_keys[j]].f.get(); } catch (const std::exception &e) { perr = e.what(); }
        c->pend_cf.erase(cf_keys[j]);
      }
    // the cache owns every awaited table from here; a failed task left none
    for (size_t j = 0; j < cheb_keys.size(); j++)
      if (queued[j] && cheb_d[j].p) c->cheb.emplace(cheb_keys[j], cheb_d[j]);
    for (size_t j = 0; j < cf_keys.size(); j++)
      if (queued[cheb_keys.size() + j] && cf_d[j].kk.p) c->cf.emplace(cf_keys[j], cf_d[j]);
    if (!perr.empty()) return fail(c, WNDGPU_E_NOMEM, "prefetched table: %s", perr.c_str());
    int r;
    for (size_t j = 0; j < cheb_keys.size(); j++) {
      if (queued[j]) continue;
      DevTable t;
      if ((r = upload(c, t, cheb_v[j].data(), cheb_v[j].size() * sizeof(double)))) return r;
      c->cheb.emplace(cheb_keys[j], t);
    }
    for (size_t j = 0; j < cf_keys.size(); j++) {
      if (queued[cheb_keys.size() + j]) continue;
      const CfTable &t = cf_v[j];
      CfDev d;
      d.np = (int)t.kk.size();
      d.NN = t.NN;
      if ((r = upload(c, d.kk, t.kk.data(), t.kk.size() * sizeof(int)))) return r;
      if ((r = upload(c, d.tn, t.tn.data(), t.tn.size() * sizeof(double)))) return r;
      if ((r = upload(c, d.fc, t.fc.data(), t.fc.size() * sizeof(double)))) return r;
      if ((r = upload(c, d.fs, t.fs.data(), t.fs.size() * sizeof(double)))) return r;
      c->cf.emplace(cf_keys[j], d);
    }
  }
  for (const auto &k : c->need_cheb) c->cheb[k].used = std::max(c->cheb[k].used, c->batch_id);
  for (const auto &k : c->need_cf) c->cf[k].used = std::max(c->cf[k].used, c->batch_id);
  c->need_cheb.clear();
  c->need_cf.clear();
  // prefetched tables no batch took: awaited and freed
  drop_pending(c, c->batch_id - 2);
  // eviction: previous batches have completed (every compute call waits for
  // its streams), so tables this batch does not read can be freed
  size_t total = 0;
  for (auto &kv : c->cheb) total += kv.second.bytes;
  for (auto &kv : c->cf) total += kv.second.bytes();
  while (total > c->table_budget) {
    long oldest = c->batch_id;
    int which = -1;
    std::pair<int, int> key;
    for (auto &kv : c->cheb)
      if (kv.second.used < oldest) { oldest = kv.second.used; which = 0; key = kv.first; }
    for (auto &kv : c->cf)
      if (kv.second.used < oldest) { oldest = kv.second.used; which = 1; key = kv.first; }
    if (which < 0) break;  // everything left is read by this batch
    drop_graphs(c);
    if (which == 0) {
      DevTable &t = c->cheb[key];
      total -= t.bytes;
      HIPCHK(c, hipFreeAsync(t.p, c->stream));
      c->cheb.erase(key);
    } else {
      CfDev &d = c->cf[key];
      total -= d.bytes();
      for (DevTable *t : {&d.kk, &d.tn, &d.fc, &d.fs}) HIPCHK(c, hipFreeAsync(t->p, c->stream));
      c->cf.erase(key);
    }
  }
  return 0;
}

static int get_dspos(wndgpu_ctx *c, int len, double dx, const double **p, int *count) {
  auto key = std::make_pair(len, dx);
  auto it = c->dspos.find(key);
  if (it == c->dspos.end()) {
    std::vector<double> v = ds_positions(len, dx);
    DevTable t;
    int r = upload(c, t, v.data(), v.size() * sizeof(double));
    if (r) return r;
    it = c->dspos.emplace(key, t).first;
  }
  *p = (const double *)it->second.p;
  *count = (int)(it->second.bytes / sizeof(double));
  return 0;
}

static int get_comb(wndgpu_ctx *c, int m, const double **p) {
  auto it = c->comb.find(m);
  if (it == c->comb.end()) {
    std::vector<double> v = comb_offsets(m);
    DevTable t;
    int r = upload(c, t, v.data(), v.size() * sizeof(double));
    if (r) return r;
    it = c->comb.emplace(m, t).first;
  }
  *p = (const double *)it->second.p;
  return 0;
}

static void out_size(int kind, int h, int w, int *oh, int *ow) {
  if (kind == T_CHEBYSHEV) { *oh = *ow = std::min(h, w); }
  else if (kind == T_WAVELET) { *oh = 2 * ((h + 9) / 2); *ow = 2 * ((w + 9) / 2); }
  else { *oh = h; *ow = w; }
}

// One launch closure of a batch: its name (timing), the kernels, whether it
// runs on the side stream, and whether it is issued after the other feature
// launches of its stream (late).
struct LaunchItem {
  std::string name;
  std::function<void(hipStream_t)> f;  // issues the launch's kernels on the given stream
  bool side;
  bool late;
};

// Root of a sampled batch: load each image's crop, Downsample it (pixels and
// colour plane), normalise it, then cut the rotated / tiled samples; the
// samples are the root entries.  A batch with one unrotated, untiled sample
// per image uses the pre-processed planes as they are.
template <class Push>
static int build_sampled_root(wndgpu_ctx *c, const wndgpu_plan *plan, const Batch &B, Arena &A, NodeRun &R,
                              Push &push) {
  hipStream_t s = c->stream;
  const int ni = (int)B.pre.size(), n = B.n;
  const bool colour = plan->color;
  std::vector<Img> cim(ni), dim(ni);
  std::vector<unsigned char *> chsv(ni, nullptr), dhsv(ni, nullptr);
  std::vector<LoadDesc> roi(ni), full(ni);
  bool wide = false;
  for (int k = 0; k < ni; k++) wide |= B.pre[k].full.dtype == 5;
  double *range = wide ? A.dalloc_n<double>(2 * (size_t)ni) : nullptr;
  long maxc = 0;
  for (int k = 0; k < ni; k++) {
    const PreImage &P = B.pre[k];
    cim[k] = Img{A.dalloc_n<double>((size_t)P.ch * P.cw), P.ch, P.cw};
    if (colour) chsv[k] = A.dalloc_n<unsigned char>(3 * (size_t)P.ch * P.cw);
    roi[k] = P.roi;
    roi[k].hsv = chsv[k];
    roi[k].range = range ? range + 2 * k : nullptr;
    full[k] = P.full;
    maxc = std::max(maxc, (long)P.ch * P.cw);
  }
  // Downsample: the pixel plane with the Chebyshev-Fourier kernels, the colour plane alongside
  std::vector<DsDesc> ds;
  std::vector<HsvDsDesc> hds;
  int mx_oh = 0, mx_ow = 0, mx_nh = 0, mx_nw = 0;
  for (int k = 0; k < ni; k++) {
    const PreImage &P = B.pre[k];
    if (!P.ds) {
      dim[k] = cim[k];
      dhsv[k] = chsv[k];
      continue;
    }
    DsDesc d;
    d.in = cim[k].p;
    d.tmp = A.dalloc_n<double>((size_t)P.ch * P.cw);
    d.out = A.dalloc_n<double>((size_t)P.dh * P.dw);
    d.oh = P.ch;
    d.ow = P.cw;
    d.nh = P.dh;
    d.nw = P.dw;
    d.dx = d.dy = P.dx;
    d.xs = d.ys = nullptr;
    d.cx = d.cy = 0;
    if (!A.sizing) {
      int r;
      if ((r = get_dspos(c, P.cw, P.dx, &d.xs, &d.cx))) return r;
      if ((r = get_dspos(c, P.ch, P.dx, &d.ys, &d.cy))) return r;
    }
    ds.push_back(d);
    dim[k] = Img{d.out, P.dh, P.dw};
    if (colour) {
      HsvDsDesc h;
      h.in = chsv[k];
      h.tmp = A.dalloc_n<unsigned char>(3 * (size_t)P.ch * P.cw);
      h.out = A.dalloc_n<unsigned char>(3 * (size_t)P.dh * P.dw);
      h.oh = d.oh; h.ow = d.ow; h.nh = d.nh; h.nw = d.nw; h.cx = d.cx; h.cy = d.cy;
      h.xs = d.xs; h.ys = d.ys; h.dx = d.dx; h.dy = d.dy;
      hds.push_back(h);
      dhsv[k] = h.out;
    }
    mx_oh = std::max(mx_oh, P.ch);
    mx_ow = std::max(mx_ow, P.cw);
    mx_nh = std::max(mx_nh, P.dh);
    mx_nw = std::max(mx_nw, P.dw);
  }
  std::vector<NormDesc> nd;
  long maxd = 0;
  for (int k = 0; k < ni; k++) {
    const PreImage &P = B.pre[k];
    if (P.norm) nd.push_back(NormDesc{dim[k].p, (long)P.dh * P.dw, P.mean, P.stdev, P.max_range, 0, 0});
    maxd = std::max(maxd, (long)P.dh * P.dw);
  }
  // samples
  const bool direct = n == ni && std::all_of(B.smp.begin(), B.smp.end(), [](const SampleRef &q) { return q.rot == 0; }) &&
                      std::all_of(B.smp.begin(), B.smp.end(), [&](const SampleRef &q) {
                        return q.x0 == 0 && q.y0 == 0 && q.h == B.pre[q.img].dh && q.w == B.pre[q.img].dw;
                      });
  std::vector<Img> sim(n);
  std::vector<SampleDesc> sd;
  R.h.resize(n);
  R.w.resize(n);
  R.hsv.assign(n, nullptr);
  long maxs = 0;
  for (int j = 0; j < n; j++) {
    const SampleRef &q = B.smp[j];
    R.h[j] = q.h;
    R.w[j] = q.w;
    maxs = std::max(maxs, (long)q.h * q.w);
    if (direct) {
      sim[j] = dim[q.img];
      R.hsv[j] = dhsv[q.img];
      continue;
    }
    sim[j] = Img{A.dalloc_n<double>((size_t)q.h * q.w), q.h, q.w};
    if (colour) R.hsv[j] = A.dalloc_n<unsigned char>(3 * (size_t)q.h * q.w);
    sd.push_back(SampleDesc{dim[q.img].p, dhsv[q.img], R.hsv[j], dim[q.img].h, dim[q.img].w, q.rot, q.x0, q.y0});
  }
  R.imgs = A.stage(sim);
  R.stats = A.dalloc_n<Stats>(n);
  R.planes.resize(n);
  for (int j = 0; j < n; j++) R.planes[j] = sim[j].p;
  if (colour) R.hsvd = A.stage(R.hsv);
  LoadDesc *rd = A.stage(roi), *fd = A.stage(full);
  Img *cd = A.stage(cim);
  DsDesc *dsd = ds.empty() ? nullptr : A.stage(ds);
  HsvDsDesc *hdd = hds.empty() ? nullptr : A.stage(hds);
  NormDesc *ndd = nd.empty() ? nullptr : A.stage(nd);
  SampleDesc *sdd = sd.empty() ? nullptr : A.stage(sd);
  Img *ri = R.imgs;
  Stats *rs = R.stats;
  const int nds = (int)ds.size(), nnd = (int)nd.size(), nsd = (int)sd.size();
  push("root", [=](hipStream_t s) {
    if (wide) launch_rgb_range(fd, range, ni, s);
    launch_load(rd, cd, ni, maxc, s);
    if (nds) launch_downsample(dsd, nds, mx_oh, mx_ow, mx_nh, mx_nw, s);
    if (hdd) launch_downsample_hsv(hdd, nds, mx_oh, mx_ow, mx_nh, mx_nw, s);
    if (nnd) launch_normalize(ndd, nnd, maxd, s);
    if (nsd) launch_sample(sdd, ri, nsd, maxs, s);
    launch_stats(ri, rs, n, s);
  });
  return 0;
}

// Build (or size) every launch of a batch; when !A.sizing the launches are
// also appended to `launches` in execution order, and *first_feature is the
// index of the first feature launch (the ones before it compute the DAG's
// transform nodes).
static int build_batch(wndgpu_ctx *c, const wndgpu_plan *plan, const Batch &B, Arena &A,
                       std::vector<LaunchItem> &launches, size_t *first_feature) {
  if (!A.sizing) {
    c->har_n = 0;
    c->cf_n = 0;
  }
  hipStream_t s = c->stream;
  const int n = B.n;
  std::map<std::vector<int>, NodeRun> nodes;
  auto maxof = [](const std::vector<int> &v) { return *std::max_element(v.begin(), v.end()); };
  auto push = [&](const std::string &name, std::function<void(hipStream_t)> f) {
    if (!A.sizing) launches.push_back(LaunchItem{name, std::move(f), false, false});
  };
  auto push_side = [&](const std::string &name, std::function<void(hipStream_t)> f) {
    if (!A.sizing) launches.push_back(LaunchItem{name, std::move(f), true, false});
  };
  auto push_late = [&](const std::string &name, std::function<void(hipStream_t)> f) {
    if (!A.sizing) launches.push_back(LaunchItem{name, std::move(f), false, true});
  };
  // ---- root: load + stats
  if (B.sampled) {
    int r = build_sampled_root(c, plan, B, A, nodes[{}], push);
    if (r) return r;
  } else {
    NodeRun &R = nodes[{}];
    R.h = B.h0;
    R.w = B.w0;
    std::vector<Img> im(n);
    long maxp = 0;
    for (int i = 0; i < n; i++) {
      im[i] = Img{A.dalloc_n<double>((size_t)R.h[i] * R.w[i]), R.h[i], R.w[i]};
      maxp = std::max(maxp, (long)R.h[i] * R.w[i]);
    }
    R.imgs = A.stage(im);
    R.stats = A.dalloc_n<Stats>(n);
    R.planes.resize(n);
    for (int i = 0; i < n; i++) R.planes[i] = im[i].p;
    // colour planes (3 bytes per pixel) when the plan reads them; wide RGB
    // samples need the per-image channel range first
    std::vector<LoadDesc> ldv = B.load;
    double *range = nullptr;
    bool wide = false;
    for (int i = 0; i < n; i++) wide |= ldv[i].dtype == 5;
    if (wide) range = A.dalloc_n<double>(2 * (size_t)n);
    R.hsv.assign(n, nullptr);
    for (int i = 0; i < n; i++) {
      if (plan->color) R.hsv[i] = A.dalloc_n<unsigned char>(3 * (size_t)R.h[i] * R.w[i]);
      ldv[i].hsv = R.hsv[i];
      ldv[i].range = range ? range + 2 * i : nullptr;
    }
    if (plan->color) R.hsvd = A.stage(R.hsv);
    LoadDesc *ld = A.stage(ldv);
    Img *ri = R.imgs;
    Stats *rs = R.stats;
    push("root", [=](hipStream_t s) {
      if (wide) launch_rgb_range(ld, range, n, s);
      launch_load(ld, ri, n, maxp, s);
      launch_stats(ri, rs, n, s);
    });
  }
  // ---- transform nodes (parents first)
  for (const auto &key : plan->nodes) {
    if (key.empty()) continue;
    std::vector<int> pk(key.begin(), key.end() - 1);
    NodeRun &P = nodes[pk];
    NodeRun &N = nodes[key];
    const int kind = key.back();
    N.chain = key;
    if (kind == T_COLOR_COPY) {  // identity: the node shares its parent's planes and stats
      N.h = P.h;
      N.w = P.w;
      N.imgs = P.imgs;
      N.stats = P.stats;
      N.planes = P.planes;
      continue;
    }
    N.h.resize(n);
    N.w.resize(n);
    std::vector<Img> im(n);
    for (int i = 0; i < n; i++) {
      out_size(kind, P.h[i], P.w[i], &N.h[i], &N.w[i]);
      im[i] = Img{A.dalloc_n<double>((size_t)N.h[i] * N.w[i]), N.h[i], N.w[i]};
    }
    N.imgs = A.stage(im);
    N.stats = A.dalloc_n<Stats>(n);
    const int mph = maxof(P.h), mpw = maxof(P.w);
    long maxp = 0;
    for (int i = 0; i < n; i++) maxp = std::max(maxp, (long)P.h[i] * P.w[i]);
    Img *pin = P.imgs, *pout = N.imgs;
    Stats *nst = N.stats;
    const std::string nm = node_text(key);
    if (kind == T_COLOR || kind == T_HUE) {  // innermost only: P is the root
      const unsigned char *const *hv = P.hsvd;
      const unsigned char *lut = (const unsigned char *)c->color_lut.p;
      const int ck = kind == T_COLOR ? 0 : 1;
      push(nm, [=](hipStream_t s) { launch_color_node(hv, lut, pout, ck, n, maxp, s); launch_stats(pout, nst, n, s); });
    } else if (kind == T_EDGE) {
      push(nm, [=](hipStream_t s) { launch_edge(pin, pout, n, maxp, s); launch_stats(pout, nst, n, s); });
    } else if (kind == T_WAVELET) {
      std::vector<double *> g(n);
      std::vector<int> gs(n);
      for (int i = 0; i < n; i++) {
        const long gsz = (long)((P.h[i] + 9) / 2) * P.w[i];
        g[i] = A.dalloc_n<double>(2 * gsz);
        gs[i] = (int)gsz;
      }
      double **gd = A.stage(g);
      int *gsd = A.stage(gs);
      push(nm, [=](hipStream_t s) { launch_wavelet(pin, gd, gsd, pout, n, mph, mpw, s); launch_stats(pout, nst, n, s); });
    } else if (kind == T_CHEBYSHEV) {
      // Chebyshev2D (chebyshev.cpp:118-165): rows against the x basis, then
      // the transposed intermediate against the y basis
      std::vector<GemmDesc> g1(n), g2(n);
      int maxN = 0, maxH = 0;
      for (int i = 0; i < n; i++) {
        const int h = P.h[i], w = P.w[i], NN = std::min(h, w);
        const double *bx = nullptr, *by = nullptr;
        if (!A.sizing) {
          int r;
          if ((r = get_cheb(c, w, NN, &bx))) return r;
          if ((r = get_cheb(c, h, NN, &by))) return r;
        } else {
          c->need_cheb.insert({w, NN});
          c->need_cheb.insert({h, NN});
        }
        double *c1 = A.dalloc_n<double>((size_t)h * NN);
        g1[i] = GemmDesc{P.planes[i], bx, c1, h, w, NN, (long)w, 1};
        g2[i] = GemmDesc{c1, by, im[i].p, NN, h, NN, 1, (long)NN};
        maxN = std::max(maxN, NN);
        maxH = std::max(maxH, h);
      }
      GemmDesc *d1 = A.stage(g1), *d2 = A.stage(g2);
      push(nm, [=](hipStream_t s) {
        launch_cheb_gemm(d1, n, maxH, maxN, s);
        launch_cheb_gemm(d2, n, maxN, maxN, s);
        launch_stats(pout, nst, n, s);
      });
    } else {  // Fourier
      std::vector<double2 *> f1(n);
      std::vector<FftPlan> pw(n), ph(n);
      for (int i = 0; i < n; i++) {
        f1[i] = A.dalloc_n<double2>((size_t)P.h[i] * P.w[i]);
        if (!A.sizing) {
          int r;
          if ((r = get_fft(c, P.w[i], &pw[i]))) return r;
          if ((r = get_fft(c, P.h[i], &ph[i]))) return r;
        }
      }
      double2 **f1d = A.stage(f1);
      FftPlan *pwd = A.stage(pw), *phd = A.stage(ph);
      push(nm, [=](hipStream_t s) { launch_fft(pin, f1d, pwd, phd, pout, n, mph, mpw, s); launch_stats(pout, nst, n, s); });
    }
    N.planes.resize(n);
    for (int i = 0; i < n; i++) N.planes[i] = im[i].p;
  }
  // ---- feature groups: one launch per algorithm over every (node, image)
  // entry that needs it, so a launch spans all DAG nodes of the batch.
  struct Entry {
    const NodeRun *N;
    int i;
    int off;
  };
  *first_feature = launches.size();
  std::vector<std::vector<Entry>> by_alg(A_COUNT);
  for (const auto &key : plan->nodes) {
    const NodeRun &N = nodes[key];
    for (const Group &g : plan->groups)
      if (g.chain == key)
        for (int i = 0; i < n; i++) by_alg[g.alg].push_back(Entry{&N, i, g.offset});
  }
  auto row_of = [&](const Entry &e) { return A.sizing ? (double *)nullptr : B.out + (size_t)e.i * B.nf + e.off; };
  // common per-launch descriptors: planes, gathered stats, output rows
  struct Launch {
    Img *imgs;
    Stats *st;
    OutRef o;
    int ne, mh, mw;
    long maxp;
    const Stats *const *stsrc;
  };
  auto make_launch = [&](const std::vector<Entry> &E) {
    Launch L;
    std::vector<Img> im(E.size());
    std::vector<const Stats *> sp(E.size());
    std::vector<double *> rows(E.size());
    L.mh = L.mw = 0;
    L.maxp = 0;
    for (size_t k = 0; k < E.size(); k++) {
      const NodeRun &N = *E[k].N;
      const int i = E[k].i;
      im[k] = Img{N.planes[i], N.h[i], N.w[i]};
      sp[k] = N.stats + i;
      rows[k] = row_of(E[k]);
      L.mh = std::max(L.mh, N.h[i]);
      L.mw = std::max(L.mw, N.w[i]);
      L.maxp = std::max(L.maxp, (long)N.h[i] * N.w[i]);
    }
    L.ne = (int)E.size();
    L.imgs = A.stage(im);
    L.stsrc = A.stage(sp);
    L.st = A.dalloc_n<Stats>(E.size());
    L.o = OutRef{A.stage(rows)};
    return L;
  };
  for (int alg = 0; alg < A_COUNT; alg++) {
    const std::vector<Entry> &E = by_alg[alg];
    if (E.empty() || alg == A_GINI) continue;
    const std::string gn = kAlgNames[alg];
    if (alg == A_PIXSTATS) continue;  // handled with Gini below
    Launch L = make_launch(E);
    Img *ni = L.imgs;
    Stats *nst = L.st;
    const Stats *const *sts = L.stsrc;
    const OutRef o = L.o;
    const int ne = L.ne, mh = L.mh, mw = L.mw;
    const long maxp = L.maxp;
    auto gather = [=](hipStream_t gs) { launch_gather_stats(sts, nst, ne, gs); };
    switch (alg) {
      case A_MSH: push(gn, [=](hipStream_t s) { gather(s); launch_msh(ni, nst, o, ne, s); }); break;
      case A_FRACTAL: push(gn, [=](hipStream_t s) { launch_fractal(ni, o, ne, mh, mw, s); }); break;
      case A_COLOR_HIST: {  // root entries only
        std::vector<unsigned char *> hv(ne);
        for (int k = 0; k < ne; k++) hv[k] = E[k].N->hsv[E[k].i];
        unsigned char **hvd = A.stage(hv);
        const unsigned char *lut = (const unsigned char *)c->color_lut.p;
        push(gn, [=](hipStream_t s) { launch_color_hist(hvd, ni, lut, o, ne, s); });
      } break;
      case A_COMB: {
        std::vector<double> iis;
        for (const Entry &e : E) {
          std::vector<double> v = comb_offsets(e.N->h[e.i]);
          iis.insert(iis.end(), v.begin(), v.end());
        }
        double *d = A.stage(iis);
        push(gn, [=](hipStream_t s) { launch_comb(ni, d, o, ne, s); });
      } break;
      case A_RADON: {
        std::vector<double *> pr(ne);
        int maxbins = 0;
        for (int k = 0; k < ne; k++) {
          const int W = E[k].N->w[E[k].i], H = E[k].N->h[E[k].i];
          const int a = W - 1 - (W - 1) / 2, b = H - 1 - (H - 1) / 2;
          const int rl = (int)ceil(sqrt((double)(a * a) + (double)(b * b))) + 1;
          pr[k] = A.dalloc_n<double>(8 * (size_t)(2 * rl + 1));  // sums, then bounds (certified form)
          maxbins = std::max(maxbins, 2 * rl + 1);
        }
        double **prd = A.stage(pr);
        const double *trig = (const double *)c->rtrig.p;
        if (!c->radon_direct && radon_certified_fits(maxbins)) {
          const double rslack = c->radon_slack;
          push(gn, [=](hipStream_t s) { launch_radon_certified(ni, trig, prd, o, ne, maxbins, rslack, s); });
        } else {
          push(gn, [=](hipStream_t s) { launch_radon(ni, trig, prd, o, ne, maxbins, mh, mw, s); });
        }
      } break;
      case A_TAMURA: {
        std::vector<double *> sc(ne);
        for (int k = 0; k < ne; k++)
          sc[k] = A.dalloc_n<double>(5 * (size_t)E[k].N->h[E[k].i] * E[k].N->w[E[k].i] + 8);
        double **scd = A.stage(sc);
        push(gn, [=](hipStream_t s) { gather(s); launch_tamura(ni, nst, scd, o, ne, mh, mw, s); });
      } break;
      case A_HARALICK: {
        std::vector<HarDesc> hd(ne);
        int *lut_all = A.dalloc_n<int>(256 * (size_t)ne);
        int *ng_all = A.dalloc_n<int>(ne);
        double *work_all = A.dalloc_n<double>(4 * (size_t)ne);
        int *order = A.dalloc_n<int>(ne);
        for (int k = 0; k < ne; k++) {
          HarDesc &H = hd[k];
          H.img = nullptr;
          H.gray = A.dalloc_n<uint8_t>((size_t)E[k].N->h[E[k].i] * E[k].N->w[E[k].i]);
          H.cooc = nullptr;
          H.P = A.dalloc_n<double>(4 * 256 * 256);
          H.Q = A.dalloc_n<double>(4 * 256 * 256);
          H.feat = A.dalloc_n<double>(4 * 14);
          H.marg = A.dalloc_n<double>(4 * 2 * 256);
          H.work = work_all + 4 * k;
          H.ng = ng_all + k;
          H.lut = lut_all + 256 * k;
          H.count = A.dalloc_n<int>(4);
        }
        if (!A.sizing) {
          c->har_ng = ng_all;
          c->har_work = work_all;
          c->har_n = ne;
        }
        HarDesc *hdd = A.stage(hd);
        push_side(gn + " [co-occurrence, f1-f13, Q]", [=](hipStream_t s2) {
          launch_gather_stats(sts, nst, ne, s2);
          launch_fill_int(lut_all, 256 * (long)ne, 0, s2);
          launch_haralick_prep(ni, nst, hdd, ne, (int)maxp, s2);
          launch_haralick_order(ng_all, ne, order, s2);
        });
        push_side(gn + " [k_har_hess]", [=](hipStream_t s2) { launch_haralick_hess(hdd, order, ne, s2); });
        push_side(gn + " [k_har_qr]", [=](hipStream_t s2) { launch_haralick_qr(hdd, order, ne, s2); });
        push_side(gn + " [mean, range]", [=](hipStream_t s2) { launch_haralick_final(hdd, o, ne, s2); });
      } break;
      case A_ZERNIKE: {
        const int nchunk = (int)std::max<long>(1, std::min<long>(64, (maxp + 4095) / 4096));
        double *part = A.dalloc_n<double>((size_t)ne * (4 + nchunk * 144));
        const double *ht = (const double *)c->zern.p;
        const int prec = plan->precision;
        push_late(gn, [=](hipStream_t s) { launch_zernike(ni, part, nchunk, ht, o, ne, s, prec); });
      } break;
      case A_CHEBYSHEV: {
        std::vector<GemmDesc> g1(ne), g2(ne);
        std::vector<Img> vals(ne);
        int maxN = 0, maxH = 0;
        for (int k = 0; k < ne; k++) {
          const int h = E[k].N->h[E[k].i], w = E[k].N->w[E[k].i], NN = std::min(20, std::min(h, w));
          const double *bx = nullptr, *by = nullptr;
          if (!A.sizing) {
            int r;
            if ((r = get_cheb(c, w, NN, &bx))) return r;
            if ((r = get_cheb(c, h, NN, &by))) return r;
          } else {
            c->need_cheb.insert({w, NN});
            c->need_cheb.insert({h, NN});
          }
          double *c1 = A.dalloc_n<double>((size_t)h * NN);
          double *c2 = A.dalloc_n<double>((size_t)NN * NN);
          g1[k] = GemmDesc{E[k].N->planes[E[k].i], bx, c1, h, w, NN, (long)w, 1};
          g2[k] = GemmDesc{c1, by, c2, NN, h, NN, 1, (long)NN};
          vals[k] = Img{c2, NN, NN};
          maxN = std::max(maxN, NN);
          maxH = std::max(maxH, h);
        }
        GemmDesc *d1 = A.stage(g1), *d2 = A.stage(g2);
        Img *vd = A.stage(vals);
        // FP64 in every precision: the order-20 contraction streams the plane
        // once at ~5 flops per byte (HBM-bound), so reduced precision buys no
        // time, and its 32-bin histogram stays exact
        push(gn, [=](hipStream_t s) {
          launch_cheb_gemm(d1, ne, maxH, maxN, s);
          launch_cheb_gemm(d2, ne, maxN, maxN, s);
          launch_hist_values(vd, 32, o, ne, s);
        });
      } break;
      case A_CHEB_FOURIER: {
        // downsample planes above 300x300 (ImageMatrix::ChebyshevFourierTransform2D)
        std::vector<DsDesc> ds;
        std::vector<CfDesc> cd(ne);
        int moh = 0, mow = 0, mnh = 0, mnw = 0;
        for (int k = 0; k < ne; k++) {
          int h = E[k].N->h[E[k].i], w = E[k].N->w[E[k].i];
          const double *plane = E[k].N->planes[E[k].i];
          if ((long)w * h > 300 * 300) {
            double rr = std::min(300.0 / (double)w, 300.0 / (double)h);
            if (rr > 1) rr = 1;
            const double dx = 1 / rr;
            const int nw = (int)(unsigned)(rr * w), nh = (int)(unsigned)(rr * h);
            DsDesc d;
            d.in = plane;
            d.tmp = A.dalloc_n<double>((size_t)h * w);
            d.out = A.dalloc_n<double>((size_t)nh * nw);
            d.oh = h; d.ow = w; d.nh = nh; d.nw = nw;
            d.dx = dx; d.dy = dx;
            d.xs = d.ys = nullptr;
            d.cx = d.cy = 0;
            if (!A.sizing) {
              int r;
              if ((r = get_dspos(c, w, dx, &d.xs, &d.cx))) return r;
              if ((r = get_dspos(c, h, dx, &d.ys, &d.cy))) return r;
            }
            ds.push_back(d);
            moh = std::max(moh, h); mow = std::max(mow, w); mnh = std::max(mnh, nh); mnw = std::max(mnw, nw);
            plane = d.out;
            h = nh;
            w = nw;
          }
          CfDesc C;
          C.img = plane;
          C.w = w;
          C.h = h;
          if (A.sizing) c->need_cf.insert({w, h});
          if (!A.sizing) {
            CfDev *t = nullptr;
            int r;
            if ((r = get_cf(c, w, h, &t))) return r;
            C.kk = (const int *)t->kk.p;
            C.tn = (const double *)t->tn.p;
            C.fc = (const double *)t->fc.p;
            C.fs = (const double *)t->fs.p;
            C.np = t->np;
            C.NN = t->NN;
          } else {
            C.kk = nullptr; C.tn = C.fc = C.fs = nullptr;
            C.np = w * h;  // upper bound for sizing
            C.NN = 23;
          }
          C.sum = A.dalloc_n<double>(2 * 23 * 23);
          cd[k] = C;
        }
        // Entries sharing one table set (same plane size) form groups of
        // CFM_GROUP for the matrix-core sums and of CF_GROUP for the
        // reference-order recomputation.  The grouping is by size, so the
        // sizing pass forms the same groups, and its pixel counts (w*h)
        // bound the real ones: nsplit and the partial buffers it sizes are
        // upper bounds.
        std::map<std::pair<int, int>, std::vector<int>> by_size;
        for (int k = 0; k < ne; k++) by_size[{cd[k].w, cd[k].h}].push_back(k);
        int nmg = 0, maxchunk = 1;
        for (auto &kv : by_size) {
          nmg += (int)((kv.second.size() + CFM_GROUP - 1) / CFM_GROUP);
          maxchunk = std::max(maxchunk, (cd[kv.second[0]].np + 63) / 64);
        }
        // about 2048 blocks for the 256 CUs, at least 8 chunks of 64 pixels per split
        const int nsplit = std::max(1, std::min(std::min(64, maxchunk / 8), (2048 + nmg - 1) / nmg));
        std::vector<CfMGroup> mgrp;
        std::vector<CfGroup> grp;
        for (auto &kv : by_size) {
          const std::vector<int> &v = kv.second;
          const CfDesc &C0 = cd[v[0]];
          for (size_t q = 0; q < v.size(); q += CFM_GROUP) {
            CfMGroup G;
            G.kk = C0.kk; G.tn = C0.tn; G.fc = C0.fc; G.fs = C0.fs;
            G.np = C0.np; G.NN = C0.NN;
            G.ne = (int)std::min<size_t>(CFM_GROUP, v.size() - q);
            for (int g = 0; g < CFM_GROUP; g++) {
              const int k = v[q + std::min(g, G.ne - 1)];
              if (g < G.ne) cd[k].part = A.dalloc_n<double>((size_t)nsplit * CFM_PART);
              G.img[g] = cd[k].img;
              G.part[g] = cd[k].part;
            }
            mgrp.push_back(G);
          }
          for (size_t q = 0; q < v.size(); q += CF_GROUP) {
            CfGroup G;
            G.kk = C0.kk; G.tn = C0.tn; G.fc = C0.fc; G.fs = C0.fs;
            G.np = C0.np; G.NN = C0.NN;
            G.ne = (int)std::min<size_t>(CF_GROUP, v.size() - q);
            for (int g = 0; g < CF_GROUP; g++) {
              const int k = v[q + std::min(g, G.ne - 1)];
              G.img[g] = cd[k].img;
              G.sum[g] = cd[k].sum;
              G.ent[g] = k;
            }
            grp.push_back(G);
          }
        }
        for (int k = 0; k < ne; k++) cd[k].nsplit = nsplit;
        int *open = A.dalloc_n<int>(ne);
        if (!A.sizing) {
          c->cf_open = open;
          c->cf_n = ne;
        }
        const int ngrp = (int)grp.size();
        CfMGroup *mgrpd = A.stage(mgrp);
        CfGroup *grpd = A.stage(grp);
        CfDesc *cdd = A.stage(cd);
        DsDesc *dsd = ds.empty() ? nullptr : A.stage(ds);
        const int nds = (int)ds.size();
        const double slack = c->cf_slack;
        // FP64 in every precision: the certified histogram needs FP64 sums
        // (an FP32 bound over the disk would leave most bins open)
        push(gn, [=](hipStream_t s) {
          if (nds) launch_downsample(dsd, nds, moh, mow, mnh, mnw, s);
          launch_cf(mgrpd, nmg, nsplit, grpd, ngrp, cdd, ne, open, slack, o, s);
        });
      } break;
      case A_EDGE: {
        std::vector<Img> md(2 * ne), mg(ne), dr(ne);
        for (int k = 0; k < ne; k++) {
          const int h = E[k].N->h[E[k].i], w = E[k].N->w[E[k].i];
          const size_t P = (size_t)h * w;
          mg[k] = Img{A.dalloc_n<double>(P), h, w};
          dr[k] = Img{A.dalloc_n<double>(P), h, w};
          md[2 * k] = mg[k];
          md[2 * k + 1] = dr[k];
        }
        std::vector<uint64_t *> sb(2 * ne);
        std::vector<long> cap(2 * ne);
        for (int k = 0; k < 2 * ne; k++) {
          cap[k] = (long)md[k].h * md[k].w;
          sb[k] = A.dalloc_n<uint64_t>(2 * (size_t)cap[k]);
        }
        Img *mgd = A.stage(mg), *drd = A.stage(dr), *mdd = A.stage(md);
        uint64_t **sbd = A.stage(sb);
        long *capd = A.stage(cap);
        Stats *st2 = A.dalloc_n<Stats>(2 * ne);
        double *med = A.dalloc_n<double>(2 * ne);
        push(gn, [=](hipStream_t s) {
          launch_prewitt(ni, mgd, drd, ne, maxp, s);
          launch_stats(mdd, st2, 2 * ne, s);
          launch_sort(mdd, sbd, capd, 2 * ne, s);
          launch_median_gini(mdd, sbd, st2, med, OutRef{nullptr}, OutRef{nullptr}, 2 * ne, s);
          launch_edge_final(mgd, drd, st2, med, o, ne, s);
        });
      } break;
      case A_OTSU:
      case A_INV_OTSU: {
        std::vector<ObjDesc> od(ne);
        for (int k = 0; k < ne; k++) {
          const size_t P = (size_t)E[k].N->h[E[k].i] * E[k].N->w[E[k].i];
          od[k].label = A.dalloc_n<int>(P);
          od[k].area = A.dalloc_n<int>(P + 1);
          od[k].xm = A.dalloc_n<long long>(P + 1);
          od[k].ym = A.dalloc_n<long long>(P + 1);
          od[k].keys = A.dalloc_n<uint64_t>(2 * (P + 1));
          od[k].dist = A.dalloc_n<double>(P + 1);
          od[k].count = A.dalloc_n<int>(1);
        }
        ObjDesc *odd = A.stage(od);
        const int inv = alg == A_INV_OTSU;
        push(gn, [=](hipStream_t s) { gather(s); launch_objects(ni, nst, inv, odd, o, ne, (int)maxp, s); });
      } break;
      case A_GABOR: {
        const int prec = plan->precision;
        const bool cert = prec || !c->gabor_direct;  // energies with bounds, counts certified
        std::vector<double *> en(ne);
        std::vector<float *> bd(cert ? ne : 0);
        for (int k = 0; k < ne; k++) {
          const size_t px = (size_t)E[k].N->h[E[k].i] * E[k].N->w[E[k].i];
          en[k] = A.dalloc_n<double>(8 * px);
          if (cert) bd[k] = A.dalloc_n<float>(8 * px);  // error bound of each energy
        }
        double **end_ = A.stage(en);
        float **bdd = cert ? A.stage(bd) : nullptr;
        double *gprm = cert ? A.dalloc_n<double>(32 * (size_t)ne) : nullptr;  // per entry: max, min, base count
        int *gflag = cert ? A.dalloc_n<int>(8 * 64 * (size_t)ne) : nullptr;  // per entry and filter: open energies
        const double *gk = (const double *)c->gabor.p;
        const void *k0 = prec == 1 ? c->gabor_f32.p : c->gabor_bhi.p, *k1 = c->gabor_bmd.p, *k2 = c->gabor_blo.p;
        const double *wsum = (const double *)c->gabor_wsum.p;
        const double gslack = c->gabor_slack;
        push_late(gn, [=](hipStream_t s) {
          if (prec) launch_gabor_mixed(prec, ni, k0, k1, k2, wsum, gk, end_, bdd, gprm, gflag, o, ne, mh, mw, s);
          else if (cert) launch_gabor_certified(ni, gk, wsum, gslack, end_, bdd, gprm, gflag, o, ne, mh, mw, s);
          else launch_gabor(ni, gk, end_, o, ne, mh, mw, s);
        });
      } break;
      default: break;
    }
  }
  // Pixel Intensity Statistics and Gini share one sort of each plane
  {
    std::map<std::pair<const NodeRun *, int>, std::pair<int, int>> ps;  // -> (pix off, gini off)
    std::vector<std::pair<const NodeRun *, int>> order;
    for (int alg : {A_PIXSTATS, A_GINI})
      for (const Entry &e : by_alg[alg]) {
        auto key = std::make_pair(e.N, e.i);
        if (!ps.count(key)) { ps[key] = {-1, -1}; order.push_back(key); }
        (alg == A_PIXSTATS ? ps[key].first : ps[key].second) = e.off;
      }
    if (!order.empty()) {
      std::vector<Entry> E;
      for (auto &k : order) E.push_back(Entry{k.first, k.second, 0});
      Launch L = make_launch(E);
      std::vector<double *> pr(E.size()), gr(E.size());
      std::vector<uint64_t *> sb(E.size());
      std::vector<long> cap(E.size());
      for (size_t k = 0; k < E.size(); k++) {
        const auto &o2 = ps[order[k]];
        pr[k] = (o2.first >= 0 && !A.sizing) ? B.out + (size_t)E[k].i * B.nf + o2.first : nullptr;
        gr[k] = (o2.second >= 0 && !A.sizing) ? B.out + (size_t)E[k].i * B.nf + o2.second : nullptr;
        cap[k] = (long)E[k].N->h[E[k].i] * E[k].N->w[E[k].i];
        sb[k] = A.dalloc_n<uint64_t>(2 * (size_t)cap[k]);
      }
      const OutRef op{A.stage(pr)}, og{A.stage(gr)};
      uint64_t **sbd = A.stage(sb);
      long *capd = A.stage(cap);
      Img *ni = L.imgs;
      Stats *nst = L.st;
      const Stats *const *sts = L.stsrc;
      const int ne = L.ne;
      push("Pixel Intensity Statistics / Gini", [=](hipStream_t s) {
        launch_gather_stats(sts, nst, ne, s);
        launch_sort(ni, sbd, capd, ne, s);
        launch_median_gini(ni, sbd, nst, nullptr, op, og, ne, s);
      });
    }
  }
  return 0;
}

static int ensure_tables(wndgpu_ctx *c) {
  if (c->gabor.p) return 0;
  int r;
  std::vector<double> g = gabor_kernels(), z = zernike_h(), t = radon_trig();
  if ((r = upload(c, c->gabor, g.data(), g.size() * sizeof(double)))) return r;
  if ((r = upload(c, c->zern, z.data(), z.size() * sizeof(double)))) return r;
  if ((r = upload(c, c->rtrig, t.data(), t.size() * sizeof(double)))) return r;
  return 0;
}

// float -> bf16 bits, round to nearest even
static unsigned short host_bf16(float f) {
  unsigned u;
  memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

// Gabor kernels for the reduced-precision forms: FP32 (re, im) as the FP64
// table; BF16x3 as hi and lo planes [j][i-block of 16][n = 2 f + part][16],
// taps 38..47 zero
static float bf16_value(unsigned short h) {
  const unsigned b = (unsigned)h << 16;
  float f;
  memcpy(&f, &b, 4);
  return f;
}

static int ensure_mixed_tables(wndgpu_ctx *c) {
  if (c->gabor_f32.p) return 0;
  const std::vector<double> g = gabor_kernels();
  const int N = 38;
  std::vector<float> f32(g.begin(), g.end());
  // w = hi + mid + lo to 2^-24 |w|, each part the bf16 nearest to the rest
  std::vector<unsigned short> hi((size_t)N * 3 * 16 * 16, 0), md(hi.size(), 0), lo(hi.size(), 0);
  for (int j = 0; j < N; j++)
    for (int ib = 0; ib < 3; ib++)
      for (int n = 0; n < 16; n++)
        for (int k = 0; k < 16; k++) {
          const int i = 16 * ib + k;
          if (i >= N) continue;
          const double w = g[(size_t)(n >> 1) * N * N * 2 + 2 * ((size_t)j * N + i) + (n & 1)];
          const unsigned short h = host_bf16((float)w);
          const double r1 = w - (double)bf16_value(h);
          const unsigned short m = host_bf16((float)r1);
          const double r2 = r1 - (double)bf16_value(m);
          const size_t o = (((size_t)(j * 3 + ib) * 16 + n) * 16) + k;
          hi[o] = h;
          md[o] = m;
          lo[o] = host_bf16((float)r2);
        }
  std::vector<double> wsum(8, 0.0);
  for (int f = 0; f < 8; f++) {
    double sm = 0.0;
    for (int e = 0; e < 2 * N * N; e++) sm += fabs(g[(size_t)f * N * N * 2 + e]);
    wsum[f] = sm * (1.0 + 1e-12);
  }
  int r;
  if ((r = upload(c, c->gabor_f32, f32.data(), f32.size() * sizeof(float)))) return r;
  if ((r = upload(c, c->gabor_bhi, hi.data(), hi.size() * 2))) return r;
  if ((r = upload(c, c->gabor_bmd, md.data(), md.size() * 2))) return r;
  if ((r = upload(c, c->gabor_blo, lo.data(), lo.size() * 2))) return r;
  if ((r = upload(c, c->gabor_wsum, wsum.data(), wsum.size() * sizeof(double)))) return r;
  return 0;
}

// the colour class of every (h, s, v) byte triple (16 MiB), computed on the
// device once per context
static int ensure_color_lut(wndgpu_ctx *c) {
  if (c->color_lut.p) return 0;
  HIPCHK(c, hipMalloc(&c->color_lut.p, (size_t)1 << 24));
  c->color_lut.bytes = (size_t)1 << 24;
  launch_color_lut((unsigned char *)c->color_lut.p, c->stream);
  HIPCHK(c, hipGetLastError());
  return 0;
}

static int issue_batch(wndgpu_ctx *c, const std::vector<LaunchItem> &launches, size_t first_feature, size_t hbytes,
                       bool timing);
static int run_batch_graph(wndgpu_ctx *c, const wndgpu_plan *plan, Batch &B, const Arena &A,
                           std::vector<LaunchItem> &launches, size_t first_feature);

static int run_batch(wndgpu_ctx *c, const wndgpu_plan *plan, Batch &B, bool timing) {
  int r;
  if ((r = ensure_tables(c))) return r;
  if (plan->color && (r = ensure_color_lut(c))) return r;
  if ((r = ensure_mixed_tables(c))) return r;  // (their filter sums bound the plain form's Gabor energies too)
  std::vector<LaunchItem> launches;
  size_t first_feature = 0;
  // pass 1: sizes
  Arena A1{c, true};
  if ((r = build_batch(c, plan, B, A1, launches, &first_feature))) return r;
  if ((r = prepare_tables(c))) return r;
  if (A1.d + 4096 > c->dev_cap || A1.h + 4096 > c->host_cap) drop_graphs(c);
  if (A1.d + 4096 > c->dev_cap) {
    if (c->dev) HIPCHK(c, hipFree(c->dev));
    c->dev = nullptr;
    size_t cap = (A1.d + 4096) + (A1.d + 4096) / 4;
    HIPCHK(c, hipMalloc(&c->dev, cap));
    c->dev_cap = cap;
  }
  if (A1.h + 4096 > c->host_cap) {
    if (c->host) HIPCHK(c, hipHostFree(c->host));
    if (c->hdesc) HIPCHK(c, hipFree(c->hdesc));
    c->host = nullptr;
    c->hdesc = nullptr;
    c->host_cap = 0;
    size_t cap = (A1.h + 4096) * 2;
    HIPCHK(c, hipHostMalloc((void **)&c->host, cap, hipHostMallocDefault));
    HIPCHK(c, hipMalloc(&c->hdesc, cap));
    c->host_cap = cap;
  }
  // pass 2: real addresses
  Arena A2{c, false};
  if ((r = build_batch(c, plan, B, A2, launches, &first_feature))) return r;
  c->times.clear();
  c->tnames.clear();
  if (c->graph_mode && !timing) return run_batch_graph(c, plan, B, A2, launches, first_feature);
  return issue_batch(c, launches, first_feature, A2.h, timing);
}

// The launch sequence of a batch: the descriptor copy, the transform nodes
// on `stream`, the side-stream launches forked after the nodes, the other
// feature launches, and the join that makes `stream` wait for the side
// stream.  Only stream operations: the sequence can be captured.
static int issue_batch(wndgpu_ctx *c, const std::vector<LaunchItem> &launches, size_t first_feature, size_t hbytes,
                       bool timing) {
  HIPCHK(c, hipMemcpyAsync(c->hdesc, c->host, hbytes, hipMemcpyHostToDevice, c->stream));
  if (timing)
    while (c->evpool.size() < 2 * launches.size()) {
      hipEvent_t e;
      HIPCHK(c, hipEventCreate(&e));
      c->evpool.push_back(e);
    }
  size_t slot = 0;
  hipError_t ev_err = hipSuccess;
  auto issue = [&](const LaunchItem &L, hipStream_t st) {
    if (timing && ev_err == hipSuccess) ev_err = hipEventRecord(c->evpool[2 * slot], st);
    L.f(st);
    if (timing) {
      if (ev_err == hipSuccess) ev_err = hipEventRecord(c->evpool[2 * slot + 1], st);
      c->tnames.push_back(L.name);
    }
    slot++;
  };
  for (size_t k = 0; k < first_feature; k++) issue(launches[k], c->stream);
  // feature launches read only the node planes and stats and write their own
  // outputs: the Haralick chain goes to the side stream and the others
  // alternate between the main and the feature stream, all forked from the
  // main stream once the nodes are issued and joined back at the end.  The
  // chain's Hessenberg and Schur kernels are memory-bound and fill every CU's
  // LDS while they run; a kernel issued meanwhile gets a CU only as their
  // blocks leave one, and holds up whatever follows it on its stream.  So
  // the long compute-bound launches (Zernike, Gabor: marked late) go last on
  // their streams, where they fill in beside the chain, after the short
  // memory-bound ones, which then run early next to the chain's first kernels.
  const bool fork = !c->serial && launches.size() > first_feature;
  if (fork) {
    HIPCHK(c, hipEventRecord(c->fork_ev, c->stream));
    HIPCHK(c, hipStreamWaitEvent(c->side, c->fork_ev, 0));
    HIPCHK(c, hipStreamWaitEvent(c->fstream, c->fork_ev, 0));
  }
  for (size_t k = first_feature; k < launches.size(); k++)
    if (launches[k].side) issue(launches[k], fork ? c->side : c->stream);
  size_t nmain = 0;
  for (const bool late : {false, true})
    for (size_t k = first_feature; k < launches.size(); k++)
      if (!launches[k].side && launches[k].late == late)
        issue(launches[k], (fork && (nmain++ & 1)) ? c->fstream : c->stream);
  if (fork) {
    HIPCHK(c, hipEventRecord(c->join_ev, c->side));
    HIPCHK(c, hipStreamWaitEvent(c->stream, c->join_ev, 0));
    HIPCHK(c, hipEventRecord(c->fjoin_ev, c->fstream));
    HIPCHK(c, hipStreamWaitEvent(c->stream, c->fjoin_ev, 0));
  }
  HIPCHK(c, ev_err);
  HIPCHK(c, hipGetLastError());
  return 0;
}

// Everything a batch's kernel arguments depend on besides the staged
// descriptors (which the captured copy refreshes on every replay): the
// plan, its precision and schedule, the arena and staging sizes and
// addresses, and the geometry of every image, pre-processing step and
// sample.  Tables are not in it: freeing one drops every cached graph.
static std::string batch_signature(const wndgpu_ctx *c, const wndgpu_plan *plan, const Batch &B, const Arena &A,
                                   size_t nlaunch) {
  std::ostringstream o;
  o << plan->uid << '|' << plan->precision << '|' << c->serial << '|' << (const void *)c->dev << '|'
    << (const void *)c->hdesc << '|' << (const void *)c->host << '|' << A.d << '|' << A.h << '|' << nlaunch << '|'
    << B.n << '|' << B.nf << '|' << B.sampled;
  for (const LoadDesc &L : B.load) o << '|' << L.h << 'x' << L.w << ',' << L.stride << ',' << L.dtype;
  for (const PreImage &P : B.pre)
    o << '|' << P.full.h << 'x' << P.full.w << ',' << P.full.stride << ',' << P.full.dtype << ';' << P.roi.h << 'x'
      << P.roi.w << ';' << P.ch << ',' << P.cw << ',' << P.dh << ',' << P.dw << ',' << P.ds << ',' << P.dx << ','
      << P.norm << ',' << P.mean << ',' << P.stdev << ',' << P.max_range;
  for (const SampleRef &q : B.smp) o << '|' << q.img << ',' << q.rot << ',' << q.x0 << ',' << q.y0 << ',' << q.h << ',' << q.w;
  return o.str();
}

// capture the launch sequence into an executable graph; on any capture
// error the stream leaves capture mode and *exec stays null
static void capture_batch(wndgpu_ctx *c, const std::vector<LaunchItem> &launches, size_t first_feature, size_t hbytes,
                          hipGraphExec_t *exec) {
  *exec = nullptr;
  if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed) != hipSuccess) return;
  const int r = issue_batch(c, launches, first_feature, hbytes, false);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (r == 0 && e == hipSuccess && g) {
    if (hipGraphInstantiate(exec, g, nullptr, nullptr, 0) != hipSuccess) *exec = nullptr;
  }
  if (g) release_check(hipGraphDestroy(g), "hipGraphDestroy");
  (void)hipGetLastError();
  if (*exec) c->graph_captures++;
}

static int run_batch_graph(wndgpu_ctx *c, const wndgpu_plan *plan, Batch &B, const Arena &A,
                           std::vector<LaunchItem> &launches, size_t first_feature) {
  const std::string key = batch_signature(c, plan, B, A, launches.size());
  auto it = c->graphs.find(key);
  if (it == c->graphs.end()) {
    // the first batch of a signature runs eagerly (tables, attributes and
    // allocations settle); the second is captured
    if (c->graph_failed.count(key) || c->graph_seen.insert(key).second)
      return issue_batch(c, launches, first_feature, A.h, false);
    wndgpu_ctx::BatchGraph G;
    const bool serial0 = c->serial;
    const int sched0 = serial0 ? 1 : 0;
    capture_batch(c, launches, first_feature, A.h, &G.exec[sched0]);
    if (c->graph_mode == 2 && G.exec[sched0]) {
      // the other schedule: the same batch built with the other stream choice
      c->serial = !serial0;
      std::vector<LaunchItem> other;
      size_t ff = 0;
      Arena A3{c, false};
      int r = build_batch(c, plan, B, A3, other, &ff);
      if (r == 0 && A3.d == A.d && A3.h == A.h) capture_batch(c, other, ff, A3.h, &G.exec[1 - sched0]);
      c->serial = serial0;
      if (r) return r;
      // restore the staging contents of the batch being run
      Arena A4{c, false};
      std::vector<LaunchItem> again;
      if ((r = build_batch(c, plan, B, A4, again, &ff))) return r;
    }
    if (!G.exec[sched0]) {
      for (hipGraphExec_t e : G.exec)
        if (e) release_check(hipGraphExecDestroy(e), "hipGraphExecDestroy");
      c->graph_failed.insert(key);
      return issue_batch(c, launches, first_feature, A.h, false);
    }
    if (c->graphs.size() >= 16) {  // keep the 16 most recently used
      auto old = c->graphs.begin();
      for (auto g = c->graphs.begin(); g != c->graphs.end(); ++g)
        if (g->second.used < old->second.used) old = g;
      for (hipGraphExec_t e : old->second.exec)
        if (e) release_check(hipGraphExecDestroy(e), "hipGraphExecDestroy");
      c->graphs.erase(old);
    }
    it = c->graphs.emplace(key, G).first;
  }
  wndgpu_ctx::BatchGraph &G = it->second;
  G.used = ++c->graph_clock;
  if (G.pending >= 0) {  // the previous call has completed (every compute call waits for its streams)
    float ms = 0;
    if (hipEventElapsedTime(&ms, c->gev0, c->gev1) == hipSuccess) G.ms[G.pending] = ms;
    G.pending = -1;
  }
  int sched = c->serial ? 1 : 0;
  if (c->graph_mode == 2 && G.exec[0] && G.exec[1]) {
    if (G.ms[0] < 0) sched = 0;
    else if (G.ms[1] < 0) sched = 1;
    else sched = G.pick = G.ms[1] < G.ms[0] ? 1 : 0;
    if (G.pick < 0) {
      if (!c->gev0) {
        HIPCHK(c, hipEventCreate(&c->gev0));
        HIPCHK(c, hipEventCreate(&c->gev1));
      }
      HIPCHK(c, hipEventRecord(c->gev0, c->stream));
      HIPCHK(c, hipGraphLaunch(G.exec[sched], c->stream));
      HIPCHK(c, hipEventRecord(c->gev1, c->stream));
      G.pending = sched;
      c->graph_replays++;
      return 0;
    }
  }
  if (!G.exec[sched]) sched = G.exec[0] ? 0 : 1;
  HIPCHK(c, hipGraphLaunch(G.exec[sched], c->stream));
  c->graph_replays++;
  return 0;
}

// ============================================================ C ABI
extern "C" {

const char *wndgpu_version(void) { return "wndgpu 0.1 (gfx950)"; }

// Scratch of wndgpu_ad_wnd_forward_noalloc: a file-scope device array, so
// that entry point needs no allocation.  Every GPU the library runs on holds
// its own copy (WNDGPU_NOALLOC_SCRATCH_BYTES) until the process ends.  The
// array's name is a host-side stand-in: host code uses only the device
// address hipGetSymbolAddress returns, looked up once per device at context
// creation with that device current.
static __device__ __align__(256) unsigned char scratch[WNDGPU_NOALLOC_SCRATCH_BYTES];
static constexpr int kMaxDevices = 64;
static void *g_scratch_dev[kMaxDevices];
static std::mutex g_scratch_mu;

static hipError_t scratch_address(int device, void **p) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void *q = nullptr;
    const hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_dev[device] = q;
  }
  *p = g_scratch_dev[device];
  return hipSuccess;
}

int wndgpu_create(int device, wndgpu_ctx **out) {
  if (!out) return WNDGPU_E_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n || device >= kMaxDevices)
    return WNDGPU_E_DEVICE;
  wndgpu_ctx *c = new wndgpu_ctx;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return WNDGPU_E_DEVICE;
  }
  if (scratch_address(device, &c->scratch) != hipSuccess) {
    release_check(hipStreamDestroy(c->stream), "hipStreamDestroy");
    delete c;
    return WNDGPU_E_DEVICE;
  }
  if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->fstream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->fjoin_ev, hipEventDisableTiming) != hipSuccess) {
    if (c->fjoin_ev) release_check(hipEventDestroy(c->fjoin_ev), "hipEventDestroy");
    if (c->join_ev) release_check(hipEventDestroy(c->join_ev), "hipEventDestroy");
    if (c->fork_ev) release_check(hipEventDestroy(c->fork_ev), "hipEventDestroy");
    if (c->fstream) release_check(hipStreamDestroy(c->fstream), "hipStreamDestroy");
    if (c->side) release_check(hipStreamDestroy(c->side), "hipStreamDestroy");
    release_check(hipStreamDestroy(c->stream), "hipStreamDestroy");
    delete c;
    return WNDGPU_E_DEVICE;
  }
  // tables are freed into the device's stream-ordered pool: keep that memory
  // in the pool for the next tables instead of returning it at every
  // synchronisation (the table budget bounds what is live)
  {
    hipMemPool_t mp;
    if (hipDeviceGetDefaultMemPool(&mp, device) == hipSuccess) {
      uint64_t keep = UINT64_MAX;
      release_check(hipMemPoolSetAttribute(mp, hipMemPoolAttrReleaseThreshold, &keep), "hipMemPoolSetAttribute");
    }
  }
  c->serial = getenv("WNDGPU_SERIAL") != nullptr;
  if (const char *cs = getenv("WNDGPU_CF_SLACK")) c->cf_slack = std::max(1.0, atof(cs));
  if (const char *gd = getenv("WNDGPU_GABOR_DIRECT")) c->gabor_direct = atoi(gd) != 0;
  if (const char *gs = getenv("WNDGPU_GABOR_SLACK")) c->gabor_slack = std::max(1.0, atof(gs));
  if (const char *rd = getenv("WNDGPU_RADON_DIRECT")) c->radon_direct = atoi(rd) != 0;
  if (const char *rs = getenv("WNDGPU_RADON_SLACK")) c->radon_slack = std::max(1.0, atof(rs));
  if (const char *gm = getenv("WNDGPU_GRAPH")) c->graph_mode = std::min(2, std::max(0, atoi(gm)));
  if (const char *mb = getenv("WNDGPU_TABLE_CACHE_MB")) c->table_budget = (size_t)std::max(0L, atol(mb)) << 20;
  // the co-occurrence kernel keeps a 128.5 KiB triangle in LDS
  *out = c;
  return WNDGPU_OK;
}

void wndgpu_destroy(wndgpu_ctx *c) {
  if (!c) return;
  release_check(hipSetDevice(c->device), "hipSetDevice");
  release_check(hipStreamSynchronize(c->side), "hipStreamSynchronize");
  release_check(hipStreamSynchronize(c->fstream), "hipStreamSynchronize");
  release_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  c->pool.reset();  // runs the queued table tasks to the end and joins the threads
  drop_pending(c, LONG_MAX);
  if (c->ustream) release_check(hipStreamDestroy(c->ustream), "hipStreamDestroy");
  drop_graphs(c);
  if (c->gev0) release_check(hipEventDestroy(c->gev0), "hipEventDestroy");
  if (c->gev1) release_check(hipEventDestroy(c->gev1), "hipEventDestroy");
  auto fr = [c](DevTable &t) {  // tables from upload(): stream-ordered pool
    if (t.p) release_check(hipFreeAsync(t.p, c->stream), "hipFreeAsync");
    t.p = nullptr;
  };
  for (auto &kv : c->cheb) fr(kv.second);
  for (auto &kv : c->fft) fr(kv.second.first);
  for (auto &kv : c->cf) { fr(kv.second.kk); fr(kv.second.tn); fr(kv.second.fc); fr(kv.second.fs); }
  for (auto &kv : c->dspos) fr(kv.second);
  for (auto &kv : c->dsdense) fr(kv.second);
  for (auto &kv : c->comb) fr(kv.second);
  fr(c->gabor); fr(c->zern); fr(c->rtrig);
  fr(c->gabor_f32); fr(c->gabor_bhi); fr(c->gabor_bmd); fr(c->gabor_blo); fr(c->gabor_wsum);
  release_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  if (c->color_lut.p) release_check(hipFree(c->color_lut.p), "hipFree");
  for (hipEvent_t e : c->evpool) release_check(hipEventDestroy(e), "hipEventDestroy");
  if (c->dev) release_check(hipFree(c->dev), "hipFree");
  if (c->hdesc) release_check(hipFree(c->hdesc), "hipFree");
  if (c->host) release_check(hipHostFree(c->host), "hipHostFree");
  release_check(hipEventDestroy(c->fork_ev), "hipEventDestroy");
  release_check(hipEventDestroy(c->join_ev), "hipEventDestroy");
  release_check(hipEventDestroy(c->fjoin_ev), "hipEventDestroy");
  release_check(hipStreamDestroy(c->fstream), "hipStreamDestroy");
  release_check(hipStreamDestroy(c->side), "hipStreamDestroy");
  release_check(hipStreamDestroy(c->stream), "hipStreamDestroy");
  delete c;
}

const char *wndgpu_last_error(const wndgpu_ctx *c) { return c ? c->err.c_str() : "null context"; }

int wndgpu_plan_standard(int set, wndgpu_plan **plan) {
  if (!plan || set < 0 || set > 3) return WNDGPU_E_ARG;
  *plan = make_standard(set);
  return WNDGPU_OK;
}

int wndgpu_plan_from_groups(const char *const *names, int ng, wndgpu_plan **plan) {
  if (!plan || !names || ng <= 0) return WNDGPU_E_ARG;
  wndgpu_plan *p = new wndgpu_plan;
  for (int i = 0; i < ng; i++) {
    int alg;
    std::vector<int> chain;
    if (!names[i] || !parse_group(names[i], &alg, &chain)) {
      delete p;
      return WNDGPU_E_ARG;
    }
    const std::string gname = std::string(kAlgNames[alg]) + " " + node_text(chain);
    for (const Group &g : p->groups)
      if (g.name == gname) { delete p; return WNDGPU_E_ARG; }  // duplicate group (Tasks.cpp:125)
    plan_add(p, alg, chain);
  }
  plan_finalize(p);
  *plan = p;
  return WNDGPU_OK;
}

void wndgpu_plan_destroy(wndgpu_plan *p) { delete p; }
int wndgpu_plan_n_features(const wndgpu_plan *p) { return p ? p->n_features : -1; }

int wndgpu_plan_set_precision(wndgpu_plan *p, int precision) {
  if (!p || precision < WNDGPU_FP64 || precision > WNDGPU_BF16X3) return WNDGPU_E_ARG;
  p->precision = precision;
  return WNDGPU_OK;
}
const char *wndgpu_plan_feature_name(const wndgpu_plan *p, int col) {
  if (!p || col < 0 || col >= p->n_features) return "";
  return p->names[col].c_str();
}
int wndgpu_plan_group_offset(const wndgpu_plan *p, const char *name) {
  if (!p || !name) return -1;
  int alg;
  std::vector<int> chain;
  if (!parse_group(name, &alg, &chain)) return -1;
  for (const Group &g : p->groups)
    if (g.alg == alg && g.chain == chain) return g.offset;
  return -1;
}

// Geometry of the samples of one image under `sp`: false (with the reason)
// if the crop or a tile falls outside the image, as the reference's
// submatrix refuses it (cmatrix.cpp:414-436).
static bool sample_geometry(const wndgpu_image &im, const wndgpu_sampling &sp, PreImage &P, std::vector<SampleRef> &out,
                            int *status) {
  P.cw = im.width;
  P.ch = im.height;
  int x = 0, y = 0;
  if (sp.roi_x >= 0) {
    if (sp.roi_y < 0 || sp.roi_w < 1 || sp.roi_h < 1 || (long)sp.roi_x + sp.roi_w > im.width ||
        (long)sp.roi_y + sp.roi_h > im.height) {
      *status = WNDGPU_IMG_BAD_SAMPLING;
      return false;
    }
    x = sp.roi_x;
    y = sp.roi_y;
    P.cw = sp.roi_w;
    P.ch = sp.roi_h;
  }
  (void)x;
  (void)y;
  P.ds = false;
  P.dx = 1;
  P.dw = P.cw;
  P.dh = P.ch;
  if (sp.downsample > 0 && sp.downsample < 100) {
    const double ratio = (double)sp.downsample / 100.0;
    const double dx = 1 / ratio;
    if (dx != 1) {
      P.ds = true;
      P.dx = dx;
      P.dw = (int)(unsigned)(ratio * P.cw);
      P.dh = (int)(unsigned)(ratio * P.ch);
    }
  }
  P.norm = sp.mean > 0;
  P.mean = sp.mean;
  P.stdev = sp.stdev;
  if (P.norm) {
    const int bits = (im.dtype == WNDGPU_U8 || im.dtype == WNDGPU_RGB8) ? 8 : (im.dtype == WNDGPU_U16 || im.dtype == WNDGPU_RGB16) ? 16 : 0;
    if (!bits) {  // the clamp range comes from the bits per sample
      *status = WNDGPU_IMG_BAD_SAMPLING;
      return false;
    }
    P.max_range = pow((double)2, bits) - 1;
  }
  const int tiles = sp.tiles_x * sp.tiles_y;
  for (int k = 0; k < sp.rotations * tiles; k++) {
    const int r = k / tiles, tx = (k % tiles) / sp.tiles_y, ty = k % sp.tiles_y;
    const int rw = (r == 1 || r == 3) ? P.dh : P.dw, rh = (r == 1 || r == 3) ? P.dw : P.dh;
    long tw = rw, th = rh;
    if (tiles != 1) {  // TrainingSet.cpp:1465-1474
      if (r == 1 || r == 3) { th = rw / sp.tiles_x; tw = rh / sp.tiles_y; }
      else { tw = rw / sp.tiles_x; th = rh / sp.tiles_y; }
    }
    const long x0 = tx * tw, y0 = ty * th;
    if (tw < 1 || th < 1 || x0 + tw > rw || y0 + th > rh) {
      *status = WNDGPU_IMG_BAD_SAMPLING;
      return false;
    }
    if (tw < WNDGPU_MIN_SIDE || th < WNDGPU_MIN_SIDE) {
      *status = WNDGPU_IMG_TOO_SMALL;
      return false;
    }
    if (tw > WNDGPU_MAX_SIDE || th > WNDGPU_MAX_SIDE) {
      *status = WNDGPU_IMG_TOO_LARGE;
      return false;
    }
    out.push_back(SampleRef{0, r, (int)x0, (int)y0, (int)th, (int)tw});
  }
  return true;
}

static int compute_common(wndgpu_ctx *c, const wndgpu_plan *plan, const wndgpu_image *images, int n, double *out,
                          int *status, bool device_io, const wndgpu_sampling *samp = nullptr) {
  if (!c || !plan || !images || n <= 0 || !out) return fail(c, WNDGPU_E_ARG, "null argument or empty batch");
  if (samp && (samp->rotations < 1 || samp->rotations > 4 || samp->tiles_x < 1 || samp->tiles_y < 1))
    return fail(c, WNDGPU_E_ARG, "sampling: rotations must be 1..4 and tiles at least 1");
  HIPCHK(c, hipSetDevice(c->device));
  Batch B;
  B.nf = plan->n_features;
  B.sampled = samp != nullptr;
  const int ns = samp ? samp->rotations * samp->tiles_x * samp->tiles_y : 1;
  std::vector<int> idx;  // accepted images
  std::vector<PreImage> pre;
  std::vector<SampleRef> smp;
  bool rejected = false;
  for (int i = 0; i < n; i++) {
    const wndgpu_image &im = images[i];
    int st = WNDGPU_IMG_OK;
    const int max_in = samp ? 32768 : WNDGPU_MAX_SIDE;  // a sampled image is cut down before the features
    if (im.dtype < 0 || im.dtype > 5 || !im.data || im.stride < im.width) st = WNDGPU_IMG_BAD_DTYPE;
    else if (plan->color && im.dtype < 4) st = WNDGPU_IMG_NOT_COLOR;
    else if (!samp && (im.height < WNDGPU_MIN_SIDE || im.width < WNDGPU_MIN_SIDE)) st = WNDGPU_IMG_TOO_SMALL;
    else if (im.height < 1 || im.width < 1) st = WNDGPU_IMG_TOO_SMALL;
    else if (im.height > max_in || im.width > max_in) st = WNDGPU_IMG_TOO_LARGE;
    if (st == WNDGPU_IMG_OK && samp) {
      PreImage P;
      std::vector<SampleRef> g;
      if (sample_geometry(im, *samp, P, g, &st)) {
        for (SampleRef &q : g) q.img = (int)pre.size();
        pre.push_back(P);
        smp.insert(smp.end(), g.begin(), g.end());
      }
    }
    if (status) status[i] = st;
    if (st == WNDGPU_IMG_OK) idx.push_back(i);
    else rejected = true;
  }
  if (idx.empty()) return fail(c, WNDGPU_E_IMAGE, "no valid image in the batch");
  B.n = (int)idx.size() * ns;
  // host inputs are staged into one device buffer
  static const size_t esz[6] = {2, 4, 8, 1, 3, 6};  // bytes per pixel
  std::vector<void *> tmp;
  size_t in_bytes = 0;
  std::vector<size_t> in_off(idx.size());
  for (size_t k = 0; k < idx.size(); k++) {
    const wndgpu_image &im = images[idx[k]];
    if (!samp) {
      B.h0.push_back(im.height);
      B.w0.push_back(im.width);
    }
    in_off[k] = in_bytes;
    in_bytes += ((size_t)im.height * im.stride * esz[im.dtype] + 255) & ~(size_t)255;
  }
  // stream-ordered staging buffers, released on every return
  struct StreamBuf {
    hipStream_t s;
    void *p = nullptr;
    ~StreamBuf() { if (p) (void)hipFreeAsync(p, s); }
  } in_buf{c->stream}, out_buf{c->stream};
  char *din = nullptr;
  double *dout = nullptr;
  if (!device_io) {
    HIPCHK(c, hipMallocAsync(&in_buf.p, in_bytes, c->stream));
    din = (char *)in_buf.p;
    for (size_t k = 0; k < idx.size(); k++) {
      const wndgpu_image &im = images[idx[k]];
      HIPCHK(c, hipMemcpyAsync(din + in_off[k], im.data, (size_t)im.height * im.stride * esz[im.dtype],
                               hipMemcpyHostToDevice, c->stream));
    }
  }
  // device outputs with every image accepted: the rows are written in place
  const bool in_place = device_io && !rejected;
  if (in_place) {
    dout = out;
  } else {
    HIPCHK(c, hipMallocAsync(&out_buf.p, (size_t)B.n * B.nf * sizeof(double), c->stream));
    dout = (double *)out_buf.p;
  }
  for (size_t k = 0; k < idx.size(); k++) {
    const wndgpu_image &im = images[idx[k]];
    LoadDesc L;
    L.src = device_io ? im.data : (const void *)(din + in_off[k]);
    L.h = im.height;
    L.w = im.width;
    L.stride = im.stride;
    L.dtype = im.dtype;
    L.hsv = nullptr;
    L.range = nullptr;
    if (!samp) {
      B.load.push_back(L);
      continue;
    }
    PreImage &P = pre[k];
    P.full = L;
    P.roi = L;
    if (samp->roi_x >= 0) {
      P.roi.src = (const char *)L.src + ((size_t)samp->roi_y * im.stride + samp->roi_x) * esz[im.dtype];
      P.roi.h = samp->roi_h;
      P.roi.w = samp->roi_w;
    }
  }
  B.pre = std::move(pre);
  B.smp = std::move(smp);
  B.out = dout;
  c->t_table_ms = 0;
  const auto t_run = std::chrono::steady_clock::now();
  int r = run_batch(c, plan, B, getenv("WNDGPU_TIMING") != nullptr);
  c->t_host_ms = ms_since(t_run);
  const auto t_dev = std::chrono::steady_clock::now();
  if (r) {
    // work already issued on any of the streams may still read the staging buffers
    (void)hipStreamSynchronize(c->side);
    (void)hipStreamSynchronize(c->fstream);
    (void)hipStreamSynchronize(c->stream);
    return r;
  }
  // scatter accepted rows: image idx[k] owns rows idx[k]*ns .. +ns-1
  const size_t rowb = (size_t)ns * B.nf;
  if (!device_io) {
    std::vector<double> h((size_t)B.n * B.nf);
    HIPCHK(c, hipMemcpyAsync(h.data(), dout, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (size_t k = 0; k < idx.size(); k++)
      memcpy(out + (size_t)idx[k] * rowb, h.data() + k * rowb, rowb * sizeof(double));
    c->t_device_ms = ms_since(t_dev);
  } else if (!in_place) {
    // one copy per run of consecutive accepted images
    for (size_t k = 0; k < idx.size();) {
      size_t e = k + 1;
      while (e < idx.size() && idx[e] == idx[e - 1] + 1) e++;
      HIPCHK(c, hipMemcpyAsync(out + (size_t)idx[k] * rowb, dout + k * rowb, (e - k) * rowb * sizeof(double),
                               hipMemcpyDeviceToDevice, c->stream));
      k = e;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (device_io) c->t_device_ms = ms_since(t_dev);
  for (size_t k = 0; k < c->tnames.size(); k++) {
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, c->evpool[2 * k], c->evpool[2 * k + 1]));
    c->times.emplace_back(c->tnames[k], ms);
  }
  c->tnames.clear();
  return rejected ? fail(c, WNDGPU_E_IMAGE, "some images were rejected (see status)") : WNDGPU_OK;
}

int wndgpu_compute(wndgpu_ctx *c, const wndgpu_plan *plan, const wndgpu_image *images, int n, double *out,
                   int *status) {
  return compute_common(c, plan, images, n, out, status, false);
}

int wndgpu_compute_device(wndgpu_ctx *c, const wndgpu_plan *plan, const wndgpu_image *images, int n, double *out,
                          int *status) {
  return compute_common(c, plan, images, n, out, status, true);
}

int wndgpu_sampling_n_samples(const wndgpu_sampling *s) {
  if (!s || s->rotations < 1 || s->rotations > 4 || s->tiles_x < 1 || s->tiles_y < 1) return -1;
  return s->rotations * s->tiles_x * s->tiles_y;
}

int wndgpu_compute_sampled(wndgpu_ctx *c, const wndgpu_plan *plan, const wndgpu_image *images, int n,
                           const wndgpu_sampling *sampling, double *out, int *status) {
  if (!sampling) return fail(c, WNDGPU_E_ARG, "null sampling");
  return compute_common(c, plan, images, n, out, status, false, sampling);
}

int wndgpu_algorithm_n_features(const char *name) {
  if (!name) return -1;
  const int a = alg_index(trim(name));
  return a < 0 ? -1 : kNFeat[a];
}

int wndgpu_algorithm(wndgpu_ctx *c, const char *name, const double *in, int h, int w, double *out, int *n_out) {
  if (!c || !name || !in || !out) return fail(c, WNDGPU_E_ARG, "null argument");
  const int a = alg_index(trim(name));
  if (a < 0) return fail(c, WNDGPU_E_ARG, "unknown feature algorithm '%s'", name);
  wndgpu_plan *p = new wndgpu_plan;
  plan_add(p, a, {});
  plan_finalize(p);
  wndgpu_image im{in, h, w, w, WNDGPU_F64};
  int st;
  int r = wndgpu_compute(c, p, &im, 1, out, &st);
  if (n_out) *n_out = p->n_features;
  delete p;
  return r;
}

struct DBuf {
  void *p = nullptr;
  ~DBuf() {
    if (p) release_check(hipFree(p), "hipFree");
  }
};
struct FreeList {
  std::vector<void *> v;
  ~FreeList() {
    for (void *p : v) release_check(hipFree(p), "hipFree");
  }
};

// One gray transform of a device plane (h x w) into dout (out_size of t):
// the single-image form of the batched node launches.
static int transform_dev(wndgpu_ctx *c, int t, double *din, int h, int w, double *dout) {
  int r;
  if ((r = ensure_tables(c))) return r;
  int th, tw;
  out_size(t, h, w, &th, &tw);
  Img ii{din, h, w}, oo{dout, th, tw};
  FreeList frees;  // every buffer below, freed on every return
  auto dm = [&](void *p, size_t bytes) {
    hipError_t e = hipMalloc((void **)p, bytes);
    if (e == hipSuccess) frees.v.push_back(*(void **)p);
    return e;
  };
  Img *dii = nullptr, *doo = nullptr;
  HIPCHK(c, dm(&dii, sizeof(Img)));
  HIPCHK(c, dm(&doo, sizeof(Img)));
  HIPCHK(c, hipMemcpy(dii, &ii, sizeof(Img), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(doo, &oo, sizeof(Img), hipMemcpyHostToDevice));
  hipStream_t s = c->stream;
  if (t == T_COLOR_COPY) {
    HIPCHK(c, hipMemcpyAsync(dout, din, (size_t)h * w * 8, hipMemcpyDeviceToDevice, s));
  } else if (t == T_EDGE) {
    launch_edge(dii, doo, 1, (long)h * w, s);
  } else if (t == T_WAVELET) {
    double *g = nullptr;
    const int gs = ((h + 9) / 2) * w;
    HIPCHK(c, dm(&g, 2 * (size_t)gs * 8));
    double **gd = nullptr;
    int *gsd = nullptr;
    HIPCHK(c, dm(&gd, sizeof(double *)));
    HIPCHK(c, dm(&gsd, sizeof(int)));
    HIPCHK(c, hipMemcpy(gd, &g, sizeof(double *), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(gsd, &gs, sizeof(int), hipMemcpyHostToDevice));
    launch_wavelet(dii, gd, gsd, doo, 1, h, w, s);
  } else if (t == T_CHEBYSHEV) {
    const int N = std::min(h, w);
    const double *bx, *by;
    if ((r = get_cheb(c, w, N, &bx))) return r;
    if ((r = get_cheb(c, h, N, &by))) return r;
    double *c1 = nullptr;
    HIPCHK(c, dm(&c1, (size_t)h * N * 8));
    GemmDesc g[2] = {GemmDesc{din, bx, c1, h, w, N, (long)w, 1}, GemmDesc{c1, by, dout, N, h, N, 1, (long)N}};
    GemmDesc *gd = nullptr;
    HIPCHK(c, dm(&gd, sizeof g));
    HIPCHK(c, hipMemcpy(gd, g, sizeof g, hipMemcpyHostToDevice));
    launch_cheb_gemm(gd, 1, h, N, s);
    launch_cheb_gemm(gd + 1, 1, N, N, s);
  } else {
    FftPlan pw, ph;
    if ((r = get_fft(c, w, &pw))) return r;
    if ((r = get_fft(c, h, &ph))) return r;
    double2 *f1 = nullptr;
    HIPCHK(c, dm(&f1, (size_t)h * w * sizeof(double2)));
    void *pd = nullptr;
    HIPCHK(c, dm(&pd, 2 * sizeof(FftPlan) + sizeof(double2 *)));
    FftPlan pls[2] = {pw, ph};
    HIPCHK(c, hipMemcpy(pd, pls, sizeof pls, hipMemcpyHostToDevice));
    double2 **f1d = (double2 **)((char *)pd + 2 * sizeof(FftPlan));
    HIPCHK(c, hipMemcpy(f1d, &f1, sizeof(double2 *), hipMemcpyHostToDevice));
    launch_fft(dii, f1d, (FftPlan *)pd, (FftPlan *)pd + 1, doo, 1, h, w, s);
  }
  HIPCHK(c, hipStreamSynchronize(s));
  HIPCHK(c, hipGetLastError());
  return WNDGPU_OK;
}

int wndgpu_transform(wndgpu_ctx *c, const char *name, const double *in, int h, int w, double *out, int *oh,
                     int *ow) {
  if (!c || !name || !in || !out) return fail(c, WNDGPU_E_ARG, "null argument");
  const int t = tr_index(trim(name));
  if (t < 0) return fail(c, WNDGPU_E_ARG, "unknown transform '%s'", name);
  if (t == T_COLOR || t == T_HUE)
    return fail(c, WNDGPU_E_ARG, "transform '%s' reads the colour plane of an RGB image (wndgpu_transform_image)", name);
  if (h < WNDGPU_MIN_SIDE || w < WNDGPU_MIN_SIDE || h > WNDGPU_MAX_SIDE || w > WNDGPU_MAX_SIDE)
    return fail(c, WNDGPU_E_ARG, "image size %dx%d out of range", w, h);
  HIPCHK(c, hipSetDevice(c->device));
  int th, tw;
  out_size(t, h, w, &th, &tw);
  DBuf din, dout;
  HIPCHK(c, hipMalloc(&din.p, (size_t)h * w * 8));
  HIPCHK(c, hipMalloc(&dout.p, (size_t)th * tw * 8));
  HIPCHK(c, hipMemcpy(din.p, in, (size_t)h * w * 8, hipMemcpyHostToDevice));
  int r = transform_dev(c, t, (double *)din.p, h, w, (double *)dout.p);
  if (r) return r;
  HIPCHK(c, hipMemcpy(out, dout.p, (size_t)th * tw * 8, hipMemcpyDeviceToHost));
  if (oh) *oh = th;
  if (ow) *ow = tw;
  return WNDGPU_OK;
}

int wndgpu_transform_image(wndgpu_ctx *c, const char *name, const wndgpu_image *im, double *out, int *oh, int *ow) {
  if (!c || !name || !im || !im->data || !out) return fail(c, WNDGPU_E_ARG, "null argument");
  const int t = tr_index(trim(name));
  if (t < 0) return fail(c, WNDGPU_E_ARG, "unknown transform '%s'", name);
  const int h = im->height, w = im->width;
  if (im->dtype < 0 || im->dtype > 5 || im->stride < w) return fail(c, WNDGPU_E_ARG, "bad sample type or stride");
  if (h < WNDGPU_MIN_SIDE || w < WNDGPU_MIN_SIDE || h > WNDGPU_MAX_SIDE || w > WNDGPU_MAX_SIDE)
    return fail(c, WNDGPU_E_ARG, "image size %dx%d out of range", w, h);
  const bool colour = t == T_COLOR || t == T_HUE;
  if (colour && im->dtype < 4)
    return fail(c, WNDGPU_E_ARG, "transform '%s' reads the colour plane: the image must be RGB", name);
  HIPCHK(c, hipSetDevice(c->device));
  int r;
  if (t == T_COLOR && (r = ensure_color_lut(c))) return r;
  static const size_t esz[6] = {2, 4, 8, 1, 3, 6};
  const size_t in_bytes = (size_t)h * im->stride * esz[im->dtype];
  int th, tw;
  out_size(t, h, w, &th, &tw);
  DBuf src, plane, hsv, range, desc, dout;
  HIPCHK(c, hipMalloc(&src.p, in_bytes));
  HIPCHK(c, hipMalloc(&plane.p, (size_t)h * w * 8));
  HIPCHK(c, hipMalloc(&dout.p, (size_t)th * tw * 8));
  HIPCHK(c, hipMalloc(&range.p, 2 * sizeof(double)));
  if (colour) HIPCHK(c, hipMalloc(&hsv.p, 3 * (size_t)h * w));
  HIPCHK(c, hipMemcpy(src.p, im->data, in_bytes, hipMemcpyHostToDevice));
  // descriptors: LoadDesc, the plane's Img, the colour plane pointer, the output Img
  struct Desc {
    LoadDesc L;
    Img in, out;
    unsigned char *hsv;
  } hd;
  hd.L = LoadDesc{src.p, h, w, im->stride, im->dtype, (unsigned char *)hsv.p, (const double *)range.p};
  hd.in = Img{(double *)plane.p, h, w};
  hd.out = Img{(double *)dout.p, th, tw};
  hd.hsv = (unsigned char *)hsv.p;
  HIPCHK(c, hipMalloc(&desc.p, sizeof hd));
  HIPCHK(c, hipMemcpy(desc.p, &hd, sizeof hd, hipMemcpyHostToDevice));
  Desc *dd = (Desc *)desc.p;
  hipStream_t s = c->stream;
  if (im->dtype == 5) launch_rgb_range(&dd->L, (double *)range.p, 1, s);
  launch_load(&dd->L, &dd->in, 1, (long)h * w, s);
  if (colour) {
    launch_color_node(&dd->hsv, (const unsigned char *)c->color_lut.p, &dd->out, t == T_COLOR ? 0 : 1, 1,
                      (long)h * w, s);
    HIPCHK(c, hipStreamSynchronize(s));
    HIPCHK(c, hipGetLastError());
  } else if ((r = transform_dev(c, t, (double *)plane.p, h, w, (double *)dout.p))) {
    return r;
  }
  HIPCHK(c, hipMemcpy(out, dout.p, (size_t)th * tw * 8, hipMemcpyDeviceToHost));
  if (oh) *oh = th;
  if (ow) *ow = tw;
  return WNDGPU_OK;
}

// ---- classifier
#define DALLOC(ctx, buf, bytes) HIPCHK(ctx, hipMalloc(&(buf).p, (bytes) ? (bytes) : 8))

int wndgpu_normalize_train(wndgpu_ctx *c, double *data, int n, int f, double *mins, double *maxs) {
  if (!c || !data || !mins || !maxs || n <= 0 || f <= 0) return fail(c, WNDGPU_E_ARG, "bad argument");
  HIPCHK(c, hipSetDevice(c->device));
  DBuf d, mn, mx;
  DALLOC(c, d, (size_t)n * f * 8);
  DALLOC(c, mn, (size_t)f * 8);
  DALLOC(c, mx, (size_t)f * 8);
  HIPCHK(c, hipMemcpy(d.p, data, (size_t)n * f * 8, hipMemcpyHostToDevice));
  launch_norm_train((double *)d.p, n, f, (double *)mn.p, (double *)mx.p, c->stream);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(data, d.p, (size_t)n * f * 8, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(mins, mn.p, (size_t)f * 8, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(maxs, mx.p, (size_t)f * 8, hipMemcpyDeviceToHost));
  return WNDGPU_OK;
}

int wndgpu_normalize_test(wndgpu_ctx *c, double *vec, int f, const double *mins, const double *maxs) {
  if (!c || !vec || !mins || !maxs || f <= 0) return fail(c, WNDGPU_E_ARG, "bad argument");
  HIPCHK(c, hipSetDevice(c->device));
  DBuf v, mn, mx;
  DALLOC(c, v, (size_t)f * 8);
  DALLOC(c, mn, (size_t)f * 8);
  DALLOC(c, mx, (size_t)f * 8);
  HIPCHK(c, hipMemcpy(v.p, vec, (size_t)f * 8, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(mn.p, mins, (size_t)f * 8, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(mx.p, maxs, (size_t)f * 8, hipMemcpyHostToDevice));
  launch_norm_test((double *)v.p, 1, f, (const double *)mn.p, (const double *)mx.p, c->stream);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(vec, v.p, (size_t)f * 8, hipMemcpyDeviceToHost));
  return WNDGPU_OK;
}

// class labels must lie in 0..C (0: a sample of no class)
static int check_labels(wndgpu_ctx *c, const int *cls, int n, int C) {
  for (int i = 0; i < n; i++)
    if (cls[i] < 0 || cls[i] > C) return fail(c, WNDGPU_E_ARG, "class label %d of sample %d outside 0..%d", cls[i], i, C);
  return 0;
}

int wndgpu_fisher(wndgpu_ctx *c, const double *data, const int *cls, int n, int f, int C, double used, double *w) {
  if (!c || !data || !cls || !w || n <= 0 || f <= 0 || C <= 0) return fail(c, WNDGPU_E_ARG, "bad argument");
  if (int r = check_labels(c, cls, n, C)) return r;
  HIPCHK(c, hipSetDevice(c->device));
  DBuf d, k, wd;
  DALLOC(c, d, (size_t)n * f * 8);
  DALLOC(c, k, (size_t)n * 4);
  DALLOC(c, wd, (size_t)f * 8);
  HIPCHK(c, hipMemcpy(d.p, data, (size_t)n * f * 8, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(k.p, cls, (size_t)n * 4, hipMemcpyHostToDevice));
  launch_fisher((const double *)d.p, (const int *)k.p, n, f, C, used, (double *)wd.p, c->stream);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(w, wd.p, (size_t)f * 8, hipMemcpyDeviceToHost));
  return WNDGPU_OK;
}

// classify2 (wnn = false) or WNNclassify (wnn = true) of n_test normalised vectors
static int classify_common(wndgpu_ctx *c, bool wnn, const double *train, const int *cls, int n, int f, int C,
                           const double *w, const double *test, int ntest, double *prob, double *norm, long *pred) {
  if (!c || !train || !cls || !w || !test || !prob || !norm || !pred || n <= 0 || f <= 0 || C <= 0 || C > 63 ||
      ntest <= 0)
    return fail(c, WNDGPU_E_ARG, "bad argument");
  if (int r = check_labels(c, cls, n, C)) return r;
  HIPCHK(c, hipSetDevice(c->device));
  DBuf tr, k, wd, te, dist, pr, nm, pd;
  DALLOC(c, tr, (size_t)n * f * 8);
  DALLOC(c, k, (size_t)n * 4);
  DALLOC(c, wd, (size_t)f * 8);
  DALLOC(c, te, (size_t)ntest * f * 8);
  DALLOC(c, dist, (size_t)ntest * n * 8);
  DALLOC(c, pr, (size_t)ntest * C * 8);
  DALLOC(c, nm, (size_t)ntest * 8);
  DALLOC(c, pd, (size_t)ntest * sizeof(long));
  HIPCHK(c, hipMemcpy(tr.p, train, (size_t)n * f * 8, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(k.p, cls, (size_t)n * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(wd.p, w, (size_t)f * 8, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(te.p, test, (size_t)ntest * f * 8, hipMemcpyHostToDevice));
  (wnn ? launch_wnn : launch_wnd5)((const double *)tr.p, (const int *)k.p, (const double *)wd.p,
                                    (const double *)te.p, n, f, C, ntest, (double *)dist.p, (double *)pr.p,
                                    (double *)nm.p, (long *)pd.p, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(prob, pr.p, (size_t)ntest * C * 8, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(norm, nm.p, (size_t)ntest * 8, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(pred, pd.p, (size_t)ntest * sizeof(long), hipMemcpyDeviceToHost));
  return WNDGPU_OK;
}

int wndgpu_wnd5(wndgpu_ctx *c, const double *train, const int *cls, int n, int f, int C, const double *w,
                const double *test, int ntest, double *prob, double *norm, long *pred) {
  return classify_common(c, false, train, cls, n, f, C, w, test, ntest, prob, norm, pred);
}

int wndgpu_wnn(wndgpu_ctx *c, const double *train, const int *cls, int n, int f, int C, const double *w,
               const double *test, int ntest, double *prob, double *norm, long *pred) {
  return classify_common(c, true, train, cls, n, f, C, w, test, ntest, prob, norm, pred);
}

// ---- the wndcharm Python package's classifier maths

int wndgpu_py_normalize(wndgpu_ctx *c, double *data, int n, int f, int flags, double *mins, double *maxs,
                        double *means, double *stdevs, int *replaced) {
  if (!c) return WNDGPU_E_ARG;
  if (!data || n <= 0 || f <= 0 || (flags & ~127)) return fail(c, WNDGPU_E_ARG, "bad argument");
  const bool zs = flags & WNDGPU_PY_ZSCORE, ronly = flags & WNDGPU_PY_REPLACE_ONLY;
  // arrays read (given) or written (computed) by the chosen form
  if (((flags & WNDGPU_PY_GIVEN_MINS) || (!zs && !ronly)) && !mins) return fail(c, WNDGPU_E_ARG, "mins is NULL");
  if (((flags & WNDGPU_PY_GIVEN_MAXS) || (!zs && !ronly)) && !maxs) return fail(c, WNDGPU_E_ARG, "maxs is NULL");
  if (zs && !ronly && (!means || !stdevs)) return fail(c, WNDGPU_E_ARG, "means / stdevs are NULL");
  HIPCHK(c, hipSetDevice(c->device));
  const size_t fb = (size_t)f * 8;
  DBuf d, a0, a1, a2, a3, fl;
  DALLOC(c, d, (size_t)n * fb);
  DALLOC(c, a0, fb);
  DALLOC(c, a1, fb);
  DALLOC(c, a2, fb);
  DALLOC(c, a3, fb);
  DALLOC(c, fl, sizeof(int));
  HIPCHK(c, hipMemcpy(d.p, data, (size_t)n * fb, hipMemcpyHostToDevice));
  double *host[4] = {mins, maxs, means, stdevs};
  void *dev[4] = {a0.p, a1.p, a2.p, a3.p};
  for (int k = 0; k < 4; k++)
    if (host[k]) HIPCHK(c, hipMemcpy(dev[k], host[k], fb, hipMemcpyHostToDevice));
  launch_py_normalize((double *)d.p, n, f, flags, (double *)a0.p, (double *)a1.p, (double *)a2.p, (double *)a3.p,
                      (int *)fl.p, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(data, d.p, (size_t)n * fb, hipMemcpyDeviceToHost));
  for (int k = 0; k < 4; k++)
    if (host[k]) HIPCHK(c, hipMemcpy(host[k], dev[k], fb, hipMemcpyDeviceToHost));
  int any = 0;
  HIPCHK(c, hipMemcpy(&any, fl.p, sizeof(int), hipMemcpyDeviceToHost));
  if (replaced) *replaced = any;
  return WNDGPU_OK;
}

int wndgpu_py_fisher(wndgpu_ctx *c, const double *data, const int *cls, int n, int f, int C, double *w) {
  if (!c) return WNDGPU_E_ARG;
  if (!data || !cls || !w || n <= 0 || f <= 0 || C <= 0) return fail(c, WNDGPU_E_ARG, "bad argument");
  if (int r = check_labels(c, cls, n, C)) return r;
  HIPCHK(c, hipSetDevice(c->device));
  DBuf d, k, wd;
  DALLOC(c, d, (size_t)n * f * 8);
  DALLOC(c, k, (size_t)n * 4);
  DALLOC(c, wd, (size_t)f * 8);
  HIPCHK(c, hipMemcpy(d.p, data, (size_t)n * f * 8, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(k.p, cls, (size_t)n * 4, hipMemcpyHostToDevice));
  launch_py_fisher((const double *)d.p, (const int *)k.p, n, f, C, (double *)wd.p, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(w, wd.p, (size_t)f * 8, hipMemcpyDeviceToHost));
  return WNDGPU_OK;
}

int wndgpu_py_wnd5(wndgpu_ctx *c, const double *train, const int *cls, int n, int f, int C, const double *w,
                   const double *test, int ntest, double *prob, double *norm, long *pred) {
  if (!c) return WNDGPU_E_ARG;
  if (!train || !cls || !w || !test || !prob || !norm || !pred || n <= 0 || f <= 0 || C <= 0 || ntest <= 0)
    return fail(c, WNDGPU_E_ARG, "bad argument");
  if (int r = check_labels(c, cls, n, C)) return r;
  // the class views of the training set: its rows grouped by class 1..C, in order
  std::vector<int> rows, start(C + 1, 0);
  for (int cl = 1; cl <= C; cl++) {
    start[cl - 1] = (int)rows.size();
    for (int i = 0; i < n; i++)
      if (cls[i] == cl) rows.push_back(i);
  }
  start[C] = (int)rows.size();
  if (rows.empty()) return fail(c, WNDGPU_E_ARG, "no training sample has a class in 1..%d", C);
  HIPCHK(c, hipSetDevice(c->device));
  DBuf tr, wd, te, rw, st, dist, sims, cm, pr, nm, pd;
  DALLOC(c, tr, (size_t)n * f * 8);
  DALLOC(c, wd, (size_t)f * 8);
  DALLOC(c, te, (size_t)ntest * f * 8);
  DALLOC(c, rw, rows.size() * 4);
  DALLOC(c, st, (size_t)(C + 1) * 4);
  DALLOC(c, dist, (size_t)ntest * n * 8);
  DALLOC(c, sims, (size_t)ntest * n * 8);
  DALLOC(c, cm, (size_t)ntest * C * 8);
  DALLOC(c, pr, (size_t)ntest * C * 8);
  DALLOC(c, nm, (size_t)ntest * 8);
  DALLOC(c, pd, (size_t)ntest * sizeof(long));
  HIPCHK(c, hipMemcpy(tr.p, train, (size_t)n * f * 8, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(wd.p, w, (size_t)f * 8, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(te.p, test, (size_t)ntest * f * 8, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(rw.p, rows.data(), rows.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(st.p, start.data(), (size_t)(C + 1) * 4, hipMemcpyHostToDevice));
  launch_py_wnd5((const double *)tr.p, (const double *)wd.p, (const double *)te.p, (const int *)rw.p,
                 (const int *)st.p, n, f, C, ntest, (double *)dist.p, (double *)sims.p, (double *)cm.p,
                 (double *)pr.p, (double *)nm.p, (long *)pd.p, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(prob, pr.p, (size_t)ntest * C * 8, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(norm, nm.p, (size_t)ntest * 8, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(pred, pd.p, (size_t)ntest * sizeof(long), hipMemcpyDeviceToHost));
  return WNDGPU_OK;
}

int wndgpu_last_haralick_work(wndgpu_ctx *c, int *tones, double *schur_flops, int cap) {
  if (!c || c->har_n == 0) return 0;
  const int k = cap < c->har_n ? cap : c->har_n;
  if (k > 0 && tones) HIPCHK(c, hipMemcpy(tones, c->har_ng, (size_t)k * sizeof(int), hipMemcpyDeviceToHost));
  if (k > 0 && schur_flops)
    HIPCHK(c, hipMemcpy(schur_flops, c->har_work, 4 * (size_t)k * sizeof(double), hipMemcpyDeviceToHost));
  return c->har_n;
}

int wndgpu_prefetch_tables(wndgpu_ctx *c, const wndgpu_plan *plan, const wndgpu_image *images, int n) {
  if (!c || !plan || !images || n <= 0) return fail(c, WNDGPU_E_ARG, "null argument or empty batch");
  HIPCHK(c, hipSetDevice(c->device));
  // the images a compute call would accept, as its sizing pass sees them
  Batch B;
  B.nf = plan->n_features;
  for (int i = 0; i < n; i++) {
    const wndgpu_image &im = images[i];
    if (im.dtype < 0 || im.dtype > 5 || !im.data || im.stride < im.width) continue;
    if (plan->color && im.dtype < 4) continue;
    if (im.height < WNDGPU_MIN_SIDE || im.width < WNDGPU_MIN_SIDE) continue;
    if (im.height > WNDGPU_MAX_SIDE || im.width > WNDGPU_MAX_SIDE) continue;
    B.h0.push_back(im.height);
    B.w0.push_back(im.width);
    LoadDesc L;
    L.src = im.data;
    L.h = im.height;
    L.w = im.width;
    L.stride = im.stride;
    L.dtype = im.dtype;
    L.hsv = nullptr;
    L.range = nullptr;
    B.load.push_back(L);
  }
  B.n = (int)B.load.size();
  if (!B.n) return WNDGPU_OK;
  int r;
  if ((r = ensure_tables(c))) return r;
  std::vector<LaunchItem> launches;
  size_t first_feature = 0;
  Arena A{c, true};
  c->need_cheb.clear();
  c->need_cf.clear();
  r = build_batch(c, plan, B, A, launches, &first_feature);
  if (!r) {
    if (!c->ustream) HIPCHK(c, hipStreamCreateWithFlags(&c->ustream, hipStreamNonBlocking));
    if (!c->pool) c->pool.reset(new TablePool(std::min(16u, std::max(1u, std::thread::hardware_concurrency()))));
    const int dev = c->device;
    hipStream_t us = c->ustream;
    // tables already cached stay until the batch after the next has run
    for (const auto &k : c->need_cheb)
      if (c->cheb.count(k)) c->cheb[k].used = std::max(c->cheb[k].used, c->batch_id + 2);
    for (const auto &k : c->need_cf)
      if (c->cf.count(k)) c->cf[k].used = std::max(c->cf[k].used, c->batch_id + 2);
    for (const auto &k : c->need_cheb)
      if (!c->cheb.count(k) && !c->pend_cheb.count(k)) {
        auto pr = std::make_shared<std::promise<DevTable>>();
        c->pend_cheb[k] = {pr->get_future(), c->batch_id};
        const int a = k.first, b = k.second;
        c->pool->push([pr, a, b, dev, us]() {
          try {
            const std::vector<double> v = cheb_basis(a, b);
            pr->set_value(upload_on(dev, us, v.data(), v.size() * sizeof(double)));
          } catch (...) {
            pr->set_exception(std::current_exception());
          }
        });
      }
    for (const auto &k : c->need_cf)
      if (!c->cf.count(k) && !c->pend_cf.count(k)) {
        auto pr = std::make_shared<std::promise<CfDev>>();
        c->pend_cf[k] = {pr->get_future(), c->batch_id};
        const int a = k.first, b = k.second;
        c->pool->push([pr, a, b, dev, us]() {
          try {
            const CfTable t = cf_table(a, b);
            pr->set_value(upload_cf_on(dev, us, t));
          } catch (...) {
            pr->set_exception(std::current_exception());
          }
        });
      }
  }
  c->need_cheb.clear();
  c->need_cf.clear();
  return r;
}

int wndgpu_last_call_times(const wndgpu_ctx *c, double *host_ms, double *table_ms, double *device_ms) {
  if (!c) return WNDGPU_E_ARG;
  if (host_ms) *host_ms = c->t_host_ms;
  if (table_ms) *table_ms = c->t_table_ms;
  if (device_ms) *device_ms = c->t_device_ms;
  return WNDGPU_OK;
}

int wndgpu_last_cf_open(wndgpu_ctx *c, int *open, int cap) {
  if (!c || c->cf_n == 0) return 0;
  const int k = cap < c->cf_n ? cap : c->cf_n;
  if (k > 0 && open) HIPCHK(c, hipMemcpy(open, c->cf_open, (size_t)k * sizeof(int), hipMemcpyDeviceToHost));
  return c->cf_n;
}

int wndgpu_last_node_times(wndgpu_ctx *c, double *ms, const char **names, int cap) {
  if (!c) return 0;
  int k = 0;
  for (auto &t : c->times) {
    if (k >= cap) break;
    if (ms) ms[k] = t.second;
    if (names) names[k] = t.first.c_str();
    k++;
  }
  return k;
}

int wndgpu_set_graph_mode(wndgpu_ctx *c, int mode) {
  if (!c || mode < 0 || mode > 2) return fail(c, WNDGPU_E_ARG, "graph mode must be 0, 1 or 2");
  if (mode == 0) drop_graphs(c);
  c->graph_mode = mode;
  return WNDGPU_OK;
}

int wndgpu_graph_stats(const wndgpu_ctx *c, long *captures, long *replays, int *cached, int *schedule_of_last) {
  if (!c) return WNDGPU_E_ARG;
  if (captures) *captures = c->graph_captures;
  if (replays) *replays = c->graph_replays;
  if (cached) *cached = (int)c->graphs.size();
  if (schedule_of_last) {
    *schedule_of_last = -1;
    long best = -1;
    for (const auto &kv : c->graphs)
      if (kv.second.used > best) { best = kv.second.used; *schedule_of_last = kv.second.pick; }
  }
  return WNDGPU_OK;
}

// ------------------------------------------------------- differentiable form
// Forward and adjoint passes of the smooth transforms and features on
// batches of device planes; see wndgpu.h.
// Chebyshev-Fourier geometry of the differentiable form: the plane the
// coefficients are taken from (downsampled above 300 x 300 pixels, as
// FeatureAlgorithms.cpp:54-69 / cmatrix.cpp:1227-1240) and NN
struct CfAdGeom {
  bool ds;
  int nh, nw, NN, nchunk;
  double dx;
};
static CfAdGeom cf_ad_geom(int h, int w) {
  CfAdGeom g{false, h, w, 0, 0, 1.0};
  if ((long)w * h > 300 * 300) {
    double rr = std::min(300.0 / (double)w, 300.0 / (double)h);
    if (rr > 1) rr = 1;
    g.ds = true;
    g.dx = 1 / rr;
    g.nw = (int)(unsigned)(rr * w);
    g.nh = (int)(unsigned)(rr * h);
  }
  int N = 11;
  const int nm = (std::min(g.nh, g.nw) - 1) / 2;
  if (N > nm) N = nm;
  g.NN = 2 * N + 1;
  g.nchunk = cf_ad_chunks(g.nh * g.nw);
  return g;
}

static int get_dsdense(wndgpu_ctx *c, int len, double dx, int nout, const double **p) {
  auto key = std::make_tuple(len, dx, nout);
  auto it = c->dsdense.find(key);
  if (it == c->dsdense.end()) {
    std::vector<double> d = ds_dense((unsigned)len, dx, (unsigned)nout);
    DevTable t;
    int r = upload(c, t, d.data(), d.size() * sizeof(double));
    if (r) return r;
    it = c->dsdense.emplace(key, t).first;
  }
  *p = (const double *)it->second.p;
  return 0;
}

// work layout of the Chebyshev-Fourier op, in doubles: tmp (n h nw, when
// downsampled), the downsampled planes (n nh nw), chunk partials, W,
// the downsampled gradient (n nh nw), then nh nw int slots
struct CfAdWork {
  size_t tmp, dsp, part, coef, gds, slot, total_bytes;
};
static CfAdWork cf_ad_work(const CfAdGeom &g, int n, int h) {
  CfAdWork L;
  const size_t dsn = g.ds ? (size_t)n * g.nh * g.nw : 0, nn2 = (size_t)g.NN * g.NN;
  size_t o = 0;
  L.tmp = o; o += g.ds ? (size_t)n * h * g.nw : 0;
  L.dsp = o; o += dsn;
  L.part = o; o += (size_t)n * g.nchunk * 2 * nn2;
  L.coef = o; o += (size_t)n * 2 * nn2;
  L.gds = o; o += dsn;
  L.slot = o;
  L.total_bytes = o * sizeof(double) + (size_t)g.nh * g.nw * sizeof(int);
  return L;
}

static int ad_dims(int op, int h, int w, int *oh, int *ow) {
  const int N = h < w ? h : w;
  switch (op) {
    case WNDGPU_AD_CHEB_FOURIER: {
      const CfAdGeom g = cf_ad_geom(h, w);
      *oh = 1;
      *ow = g.NN * g.NN;
      return g.NN >= 1 ? 0 : -1;
    }
    case WNDGPU_AD_FOURIER: *oh = h; *ow = w; return 0;
    case WNDGPU_AD_CHEBYSHEV: *oh = N; *ow = N; return 0;
    case WNDGPU_AD_CHEBYSHEV_COEFFS: *oh = N < 20 ? N : 20; *ow = *oh; return 0;
    case WNDGPU_AD_ZERNIKE: *oh = 1; *ow = 72; return 0;
    case WNDGPU_AD_COMB: *oh = 1; *ow = 320; return 0;
    case WNDGPU_AD_GABOR: *oh = 8 * h; *ow = w; return 0;
    default: return -1;
  }
}

int wndgpu_ad_output_shape(int op, int h, int w, int *out_h, int *out_w) {
  if (!out_h || !out_w || h < WNDGPU_MIN_SIDE || w < WNDGPU_MIN_SIDE || h > WNDGPU_MAX_SIDE || w > WNDGPU_MAX_SIDE)
    return WNDGPU_E_ARG;
  return ad_dims(op, h, w, out_h, out_w) ? WNDGPU_E_ARG : WNDGPU_OK;
}

size_t wndgpu_ad_work_bytes(int op, int n, int h, int w) {
  int oh, ow;
  if (n <= 0 || ad_dims(op, h, w, &oh, &ow)) return 0;
  switch (op) {
    case WNDGPU_AD_FOURIER: return (size_t)n * h * w * sizeof(double2);
    case WNDGPU_AD_CHEB_FOURIER: return cf_ad_work(cf_ad_geom(h, w), n, h).total_bytes;
    case WNDGPU_AD_COMB: return (size_t)n * 80 * 5 * sizeof(double);
    case WNDGPU_AD_GABOR: return (size_t)n * h * w * 2 * sizeof(double);
    case WNDGPU_AD_CHEBYSHEV:
    case WNDGPU_AD_CHEBYSHEV_COEFFS: return (size_t)n * h * oh * sizeof(double);
    default: return zernike_ad_work_doubles(n, h, w) * sizeof(double);
  }
}

static int ad_check(wndgpu_ctx *c, int op, const void *in, const void *out, int n, int h, int w, const void *work) {
  int oh, ow;
  if (!in || !out || !work || n <= 0) return fail(c, WNDGPU_E_ARG, "null argument or empty batch");
  if (h < WNDGPU_MIN_SIDE || w < WNDGPU_MIN_SIDE || h > WNDGPU_MAX_SIDE || w > WNDGPU_MAX_SIDE)
    return fail(c, WNDGPU_E_ARG, "image size %dx%d out of range", w, h);
  if (ad_dims(op, h, w, &oh, &ow)) return fail(c, WNDGPU_E_ARG, "unknown differentiable operation %d", op);
  return 0;
}

// Chebyshev transform of order N as two strided products (forward) or
// their transposes (adjoint); bx [w][N], by [h][N] are the host-built bases.
static int ad_cheb_tables(wndgpu_ctx *c, int h, int w, int N, const double **bx, const double **by) {
  int r;
  if ((r = get_cheb(c, w, N, bx))) return r;
  return get_cheb(c, h, N, by);
}

// Downsample (when the plane is above 300 x 300) into the work buffer and
// fetch the coefficient tables of the resulting size
static int ad_cf_prepare(wndgpu_ctx *c, const double *in, int n, int h, int w, double *wk, hipStream_t s,
                         const double **plane, const CfDev **tab, CfAdGeom *geom) {
  const CfAdGeom g = cf_ad_geom(h, w);
  const CfAdWork L = cf_ad_work(g, n, h);
  int r;
  *plane = in;
  if (g.ds) {
    // tmp[y][k] = sum_x in[y][x] Dx[k][x];  planes[j][k] = sum_y Dy[j][y] tmp[y][k]
    const double *Dx, *Dy;
    if ((r = get_dsdense(c, w, g.dx, g.nw, &Dx)) || (r = get_dsdense(c, h, g.dx, g.nh, &Dy))) return r;
    double *tmp = wk + L.tmp, *dsp = wk + L.dsp;
    launch_ad_gemm(AdGemm{in, (long)h * w, w, 1, Dx, 0, 1, w, tmp, (long)h * g.nw, h, w, g.nw, 1.0}, n, s);
    launch_ad_gemm(AdGemm{Dy, 0, h, 1, tmp, (long)h * g.nw, g.nw, 1, dsp, (long)g.nh * g.nw, g.nh, h, g.nw, 1.0}, n,
                   s);
    *plane = dsp;
  }
  CfDev *t = nullptr;
  if ((r = get_cf(c, g.nw, g.nh, &t))) return r;
  if (t->NN != g.NN) return fail(c, WNDGPU_E_ARG, "Chebyshev-Fourier order mismatch");
  *tab = t;
  *geom = g;
  return 0;
}

int wndgpu_ad_forward(wndgpu_ctx *c, int op, const double *in, int n, int h, int w, double *out, void *work,
                      void *stream) {
  if (!c) return WNDGPU_E_ARG;
  int r = ad_check(c, op, in, out, n, h, w, work);
  if (r) return r;
  HIPCHK(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream;
  int oh, ow;
  ad_dims(op, h, w, &oh, &ow);
  if (op == WNDGPU_AD_FOURIER) {
    FftPlan pw, ph;
    if ((r = get_fft(c, w, &pw)) || (r = get_fft(c, h, &ph))) return r;
    launch_ad_fourier_fwd(in, out, (double2 *)work, n, h, w, pw, ph, s);
  } else if (op == WNDGPU_AD_ZERNIKE) {
    launch_ad_zernike_fwd(in, out, (double *)work, n, h, w, s);
  } else if (op == WNDGPU_AD_COMB) {
    launch_ad_comb_fwd(in, out, n, h, w, s);
  } else if (op == WNDGPU_AD_GABOR) {
    if ((r = ensure_tables(c))) return r;
    launch_ad_gabor_fwd(in, (const double *)c->gabor.p, out, n, h, w, s);
  } else if (op == WNDGPU_AD_CHEB_FOURIER) {
    const double *plane;
    const CfDev *t;
    CfAdGeom g;
    if ((r = ad_cf_prepare(c, in, n, h, w, (double *)work, s, &plane, &t, &g))) return r;
    const CfAdWork L = cf_ad_work(g, n, h);
    launch_ad_cf_fwd(plane, n, (long)g.nh * g.nw, (const int *)t->kk.p, (const double *)t->tn.p,
                     (const double *)t->fc.p, (const double *)t->fs.p, t->np, t->NN, g.nchunk,
                     (double *)work + L.part, out, s);
  } else {
    const int N = oh;
    const double *bx, *by;
    if ((r = ad_cheb_tables(c, h, w, N, &bx, &by))) return r;
    double *t1 = (double *)work;
    launch_ad_gemm(AdGemm{in, (long)h * w, w, 1, bx, 0, N, 1, t1, (long)h * N, h, w, N, 0.5}, n, s);
    launch_ad_gemm(AdGemm{t1, (long)h * N, 1, N, by, 0, N, 1, out, (long)N * N, N, h, N, 0.5}, n, s);
  }
  HIPCHK(c, hipGetLastError());
  return WNDGPU_OK;
}

int wndgpu_ad_backward(wndgpu_ctx *c, int op, const double *in, const double *grad_out, int n, int h, int w,
                       double *grad_in, void *work, void *stream) {
  if (!c) return WNDGPU_E_ARG;
  int r = ad_check(c, op, in, grad_in, n, h, w, work);
  if (r) return r;
  if (!grad_out) return fail(c, WNDGPU_E_ARG, "null argument");
  HIPCHK(c, hipSetDevice(c->device));
  hipStream_t s = (hipStream_t)stream;
  int oh, ow;
  ad_dims(op, h, w, &oh, &ow);
  if (op == WNDGPU_AD_FOURIER) {
    FftPlan pw, ph;
    if ((r = get_fft(c, w, &pw)) || (r = get_fft(c, h, &ph))) return r;
    launch_ad_fourier_bwd(in, grad_out, grad_in, (double2 *)work, n, h, w, pw, ph, s);
  } else if (op == WNDGPU_AD_ZERNIKE) {
    launch_ad_zernike_bwd(in, grad_out, grad_in, (double *)work, n, h, w, s);
  } else if (op == WNDGPU_AD_COMB) {
    launch_ad_comb_bwd(in, grad_out, grad_in, (double *)work, n, h, w, s);
  } else if (op == WNDGPU_AD_GABOR) {
    if ((r = ensure_tables(c))) return r;
    launch_ad_gabor_bwd(in, (const double *)c->gabor.p, grad_out, grad_in, (double *)work, n, h, w, s);
  } else if (op == WNDGPU_AD_CHEB_FOURIER) {
    const double *plane;
    const CfDev *t;
    CfAdGeom g;
    if ((r = ad_cf_prepare(c, in, n, h, w, (double *)work, s, &plane, &t, &g))) return r;
    const CfAdWork L = cf_ad_work(g, n, h);
    double *wk = (double *)work;
    int *slot = (int *)(wk + L.slot);
    launch_ad_cf_slots((const int *)t->kk.p, t->np, (long)g.nh * g.nw, slot, s);
    double *gplane = g.ds ? wk + L.gds : grad_in;
    launch_ad_cf_bwd(plane, grad_out, n, (long)g.nh * g.nw, (const int *)t->kk.p, slot, (const double *)t->tn.p,
                     (const double *)t->fc.p, (const double *)t->fs.p, t->np, t->NN, g.nchunk, wk + L.part,
                     wk + L.coef, gplane, s);
    if (g.ds) {
      // U[y][k] = sum_j Dy[j][y] G[j][k];  grad[y][x] = sum_k U[y][k] Dx[k][x]
      const double *Dx, *Dy;
      if ((r = get_dsdense(c, w, g.dx, g.nw, &Dx)) || (r = get_dsdense(c, h, g.dx, g.nh, &Dy))) return r;
      double *u = wk + L.tmp;
      launch_ad_gemm(AdGemm{Dy, 0, 1, h, gplane, (long)g.nh * g.nw, g.nw, 1, u, (long)h * g.nw, h, g.nh, g.nw, 1.0},
                     n, s);
      launch_ad_gemm(AdGemm{u, (long)h * g.nw, g.nw, 1, Dx, 0, w, 1, grad_in, (long)h * w, h, g.nw, w, 1.0}, n, s);
    }
  } else {
    const int N = oh;
    const double *bx, *by;
    if ((r = ad_cheb_tables(c, h, w, N, &bx, &by))) return r;
    double *u = (double *)work;
    // U[y][j] = 1/2 sum_k by[y][k] G[j][k];  grad[y][x] = 1/2 sum_j U[y][j] bx[x][j]
    launch_ad_gemm(AdGemm{by, 0, N, 1, grad_out, (long)N * N, 1, N, u, (long)h * N, h, N, N, 0.5}, n, s);
    launch_ad_gemm(AdGemm{u, (long)h * N, N, 1, bx, 0, 1, N, grad_in, (long)h * w, h, N, w, 0.5}, n, s);
  }
  HIPCHK(c, hipGetLastError());
  return WNDGPU_OK;
}

// WND work layout: v [nt][F], dist [nt][T], prob [nt][C], norm [nt], pred [nt], gsn [nt][C+1], gd [nt][T]
size_t wndgpu_ad_wnd_work_bytes(int n_samples, int n_features, int n_classes, int n_test) {
  if (n_samples <= 0 || n_features <= 0 || n_classes <= 0 || n_test <= 0) return 0;
  const size_t nt = n_test;
  return 8 * (nt * n_features + 2 * nt * n_samples + nt * n_classes + 2 * nt + nt * (n_classes + 1));
}

static int ad_wnd_common(wndgpu_ctx *c, const double *train, const int *cls, int n, int f, int C, const double *wts,
                         const double *mins, const double *maxs, const double *x, int nt, void *work, hipStream_t s,
                         double *prob_out) {
  if (!train || !cls || !wts || !mins || !maxs || !x || !work || n <= 0 || f <= 0 || C <= 0 || C > 63 || nt <= 0)
    return fail(c, WNDGPU_E_ARG, "bad argument");
  HIPCHK(c, hipSetDevice(c->device));
  double *v = (double *)work, *dist = v + (size_t)nt * f, *prob = dist + (size_t)nt * n, *norm = prob + (size_t)nt * C;
  long *pred = (long *)(norm + nt);
  HIPCHK(c, hipMemcpyAsync(v, x, (size_t)nt * f * 8, hipMemcpyDeviceToDevice, s));
  launch_norm_test(v, nt, f, mins, maxs, s);
  launch_wnd5(train, cls, wts, v, n, f, C, nt, dist, prob, norm, pred, s);
  if (prob_out) HIPCHK(c, hipMemcpyAsync(prob_out, prob, (size_t)nt * C * 8, hipMemcpyDeviceToDevice, s));
  return 0;
}

int wndgpu_ad_wnd_forward(wndgpu_ctx *c, const double *train, const int *classes, int n_samples, int n_features,
                          int n_classes, const double *weights, const double *mins, const double *maxs,
                          const double *x, int n_test, double *prob, void *work, void *stream) {
  if (!c) return WNDGPU_E_ARG;
  if (!prob) return fail(c, WNDGPU_E_ARG, "null argument");
  hipStream_t s = (hipStream_t)stream;
  int r = ad_wnd_common(c, train, classes, n_samples, n_features, n_classes, weights, mins, maxs, x, n_test, work, s,
                        prob);
  if (r) return r;
  HIPCHK(c, hipGetLastError());
  return WNDGPU_OK;
}

int wndgpu_ad_wnd_forward_noalloc(wndgpu_ctx *c, const double *train, const int *classes, int n_samples,
                                  int n_features, int n_classes, const double *weights, const double *mins,
                                  const double *maxs, const double *x, int n_test, double *prob, void *stream) {
  if (!c) return WNDGPU_E_ARG;
  if (!prob) return fail(c, WNDGPU_E_ARG, "null argument");
  if (n_samples <= 0 || n_features <= 0 || n_classes <= 0 || n_test <= 0)
    return fail(c, WNDGPU_E_ARG, "bad argument");
  // the forward pass uses the normalised copy, distances, probabilities,
  // norm factors and predictions of the work layout (ad_wnd_common); every
  // element is written before it is read, so nothing is cleared
  const double need = 8.0 * (double)n_test * ((double)n_features + (double)n_samples + (double)n_classes + 2.0);
  if (need > (double)WNDGPU_NOALLOC_SCRATCH_BYTES)
    return fail(c, WNDGPU_E_ARG,
                "wndgpu_ad_wnd_forward_noalloc: %d test vectors against %d samples of %d features need %.0f bytes "
                "of scratch, above the %u-byte maximum; split the test vectors into smaller calls",
                n_test, n_samples, n_features, need, (unsigned)WNDGPU_NOALLOC_SCRATCH_BYTES);
  if (!c->scratch) return fail(c, WNDGPU_E_DEVICE, "no scratch address for device %d", c->device);
  hipStream_t s = (hipStream_t)stream;
  int r = ad_wnd_common(c, train, classes, n_samples, n_features, n_classes, weights, mins, maxs, x, n_test,
                        c->scratch, s, prob);
  if (r) return r;
  HIPCHK(c, hipGetLastError());
  return WNDGPU_OK;
}

int wndgpu_ad_wnd_backward(wndgpu_ctx *c, const double *train, const int *classes, int n_samples, int n_features,
                           int n_classes, const double *weights, const double *mins, const double *maxs,
                           const double *x, int n_test, const double *grad_prob, double *grad_x, void *work,
                           void *stream) {
  if (!c) return WNDGPU_E_ARG;
  if (!grad_prob || !grad_x) return fail(c, WNDGPU_E_ARG, "null argument");
  hipStream_t s = (hipStream_t)stream;
  int r = ad_wnd_common(c, train, classes, n_samples, n_features, n_classes, weights, mins, maxs, x, n_test, work, s,
                        nullptr);
  if (r) return r;
  const size_t nt = n_test, f = n_features, n = n_samples, C = n_classes;
  double *v = (double *)work, *dist = v + nt * f, *prob = dist + nt * n, *norm = prob + nt * C;
  double *gsn = norm + 2 * nt, *gd = gsn + nt * (C + 1);
  launch_ad_wnd_bwd(train, classes, weights, x, v, mins, maxs, dist, prob, norm, grad_prob, gsn, gd, n_samples,
                    n_features, n_classes, n_test, grad_x, s);
  HIPCHK(c, hipGetLastError());
  return WNDGPU_OK;
}

}  // extern "C"
