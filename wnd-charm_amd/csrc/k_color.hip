// Colour images: the fuzzy colour class table, the "Color" and "Hue"
// transforms (src/cmatrix.cpp:1023-1066) and the Color Histogram feature
// (src/FeatureAlgorithms.cpp:502-531).  The colour plane itself (H, S, V
// bytes per pixel) is written by k_load from the RGB input.
#include "color.hpp"
#include "kernels.hpp"

namespace wnd {

// the class of every (h, s, v) byte triple: 256^3 bytes, once per context
__global__ void __launch_bounds__(256) k_color_lut(unsigned char *lut) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;  // grid covers 256^3 exactly
  lut[e] = (unsigned char)fuzzy_class(e >> 16, (e >> 8) & 255, e & 255);
}

// "Color": int(255 * class / 19) (cmatrix.cpp:1023-1047); "Hue": the H byte
// (cmatrix.cpp:1050-1066).  One thread per pixel; kind 0 = Color, 1 = Hue.
__global__ void k_color_node(const unsigned char *const *hsv, const unsigned char *lut, const Img *out, int kind) {
  const Img o = out[blockIdx.y];
  const long n = (long)o.h * o.w;
  const gptr<const unsigned char> c = to_global(hsv[blockIdx.y]);
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const unsigned h = c[3 * i], s = c[3 * i + 1], v = c[3 * i + 2];
    if (kind == 0) {
      const unsigned long idx = lut[(h << 16) | (s << 8) | v];
      const double max_range = 255.0;  // pow(2, 8) - 1
      o.p[i] = (double)int((max_range * idx) / kColorClasses);
    } else {
      o.p[i] = (double)h;
    }
  }
}

// Color Histogram: counts of the 20 class slots (slot 0 stays empty) / P.
// One block per image.
__global__ void __launch_bounds__(256) k_color_hist(const unsigned char *const *hsv, const Img *imgs,
                                                    const unsigned char *lut, OutRef o) {
  __shared__ unsigned int cnt[kColorClasses + 1];
  const Img im = imgs[blockIdx.x];
  const long n = (long)im.h * im.w;
  const gptr<const unsigned char> c = to_global(hsv[blockIdx.x]);
  if (threadIdx.x <= kColorClasses) cnt[threadIdx.x] = 0;
  __syncthreads();
  for (long i = threadIdx.x; i < n; i += blockDim.x) {
    const unsigned h = c[3 * i], s = c[3 * i + 1], v = c[3 * i + 2];
    atomicAdd(&cnt[lut[(h << 16) | (s << 8) | v]], 1u);
  }
  __syncthreads();
  if (threadIdx.x <= kColorClasses) {
    const unsigned wh = (unsigned)im.w * (unsigned)im.h;
    o.rows[blockIdx.x][threadIdx.x] = (double)cnt[threadIdx.x] / wh;
  }
}

// Wide RGB samples are scaled to bytes by the global min / max over the three
// channels of the image (LoadTIFF, cmatrix.cpp:167-193): range[2*i] = min,
// range[2*i+1] = max.  One block per image.
__global__ void __launch_bounds__(256) k_rgb_range(const LoadDesc *d, double *range) {
  __shared__ double smin[256], smax[256];
  const LoadDesc L = d[blockIdx.x];
  double mn = 1e308, mx = -1e308;
  if (L.dtype == 5) {
    const gptr<const uint16_t> p = to_global((const uint16_t *)L.src);
    for (long i = threadIdx.x; i < (long)L.h * L.w; i += blockDim.x) {
      const int y = i / L.w, x = i - (long)y * L.w;
      const long e = 3 * ((long)y * L.stride + x);
      for (int ch = 0; ch < 3; ch++) {
        const double v = p[e + ch];
        mn = fmin(mn, v);
        mx = fmax(mx, v);
      }
    }
  }
  smin[threadIdx.x] = mn;
  smax[threadIdx.x] = mx;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) {
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + k]);
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + k]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    range[2 * blockIdx.x] = smin[0];
    range[2 * blockIdx.x + 1] = smax[0];
  }
}

void launch_color_lut(unsigned char *lut, hipStream_t s) { k_color_lut<<<(1 << 24) / 256, 256, 0, s>>>(lut); }
void launch_color_node(const unsigned char *const *hsv, const unsigned char *lut, const Img *out, int kind, int nimg,
                       long maxp, hipStream_t s) {
  long blocks = (maxp + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  k_color_node<<<dim3(blocks, nimg), 256, 0, s>>>(hsv, lut, out, kind);
}
void launch_color_hist(const unsigned char *const *hsv, const Img *imgs, const unsigned char *lut, OutRef o, int nimg,
                       hipStream_t s) {
  k_color_hist<<<nimg, 256, 0, s>>>(hsv, imgs, lut, o);
}
void launch_rgb_range(const LoadDesc *d, double *range, int nimg, hipStream_t s) {
  k_rgb_range<<<nimg, 256, 0, s>>>(d, range);
}

}  // namespace wnd
