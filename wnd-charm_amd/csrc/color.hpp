// Colour pixels on the GPU: the reference's colour plane (RGB2HSV,
// RGB2GRAY, src/cmatrix.h:95-160) and its fuzzy colour classes (FindColor,
// src/colors/FuzzyCalc.cpp:594-831).  A colour class is the argmax over the
// 19 classes of the sum, over the class's rules in table order, of
// hue membership x saturation membership x value membership.
#pragma once
#include "common.hpp"

namespace wnd {

struct HSVb {
  unsigned char h, s, v;
};

// RGB2HSV (cmatrix.h:95-126): bytes of H (0..239), S and V (0..240)
__device__ __forceinline__ HSVb rgb_to_hsv(unsigned char R, unsigned char G, unsigned char B) {
  const double r = (double)R / 255, g = (double)G / 255, b = (double)B / 255;
  const double mx = fmax(r, fmax(g, b)), mn = fmin(r, fmin(g, b));
  const double delta = mx - mn;
  HSVb o;
  o.v = (unsigned char)(mx * 240.0);
  o.s = mx != 0.0 ? (unsigned char)((delta / mx) * 240.0) : 0;
  if (o.s == 0) {
    o.h = 0;
  } else {
    double h = 0;
    if (r == mx) h = (g - b) / delta;
    else if (g == mx) h = 2 + (b - r) / delta;
    else if (b == mx) h = 4 + (r - g) / delta;
    h *= 60.0;
    if (h >= 360) h -= 360.0;
    if (h < 0.0) h += 360.0;
    o.h = (unsigned char)(h * (240.0 / 360.0));
  }
  return o;
}
// RGB2GRAY (cmatrix.h:158-160)
__device__ __forceinline__ double rgb_gray(unsigned char R, unsigned char G, unsigned char B) {
  return 0.2989 * R + 0.5870 * G + 0.1140 * B;
}

constexpr int kColorClasses = 19;  // white .. light fucia = 1 .. 19 (FuzzyCalc.h:113-131)
// hue trapezoids (class, start, plateau start, plateau end, end) in the
// order of the rules text's color_functions section; red has two
struct HueTrapezoid {
  int cls;
  double a, b, c, d;
};
constexpr HueTrapezoid kHueTraps[11] = {{5, 0, 0, 0, 11},       {9, 0, 17, 17, 27},     {10, 20, 27, 27, 37},
                                        {11, 27, 39, 39, 47},   {13, 40, 50, 50, 80},   {14, 50, 80, 80, 120},
                                        {16, 80, 120, 120, 160}, {17, 120, 160, 160, 200}, {18, 160, 200, 200, 220},
                                        {19, 200, 220, 220, 230}, {5, 220, 240, 240, 240}};
// hue classes of the rule blocks, in table order
constexpr int kRuleHue[10] = {5, 9, 10, 11, 13, 14, 16, 17, 18, 19};
// rule output class for (hue block, saturation 1..6, value 1..5) in table
// order; 0 marks the two rules whose colour name the reference does not
// recognise, which never count
constexpr unsigned char kRuleClass[10][6][5] = {
    {{4, 4, 3, 2, 1}, {4, 3, 3, 19, 6}, {4, 7, 18, 6, 6}, {4, 7, 8, 19, 6}, {4, 7, 7, 5, 5}, {4, 4, 5, 5, 5}},
    {{4, 3, 3, 2, 1}, {4, 3, 3, 2, 1}, {4, 7, 3, 8, 6}, {4, 7, 7, 8, 10}, {4, 7, 7, 8, 9}, {4, 7, 7, 8, 9}},
    {{4, 3, 3, 2, 1}, {4, 3, 3, 2, 1}, {4, 3, 3, 2, 10}, {4, 7, 8, 8, 10}, {4, 7, 8, 8, 10}, {4, 7, 8, 8, 10}},
    {{4, 3, 3, 2, 1}, {4, 3, 3, 2, 1}, {4, 3, 12, 12, 11}, {4, 12, 12, 12, 11}, {4, 12, 12, 12, 11}, {4, 12, 12, 12, 11}},
    {{4, 3, 3, 2, 1}, {4, 3, 3, 2, 1}, {4, 12, 12, 13, 13}, {4, 14, 14, 13, 13}, {4, 14, 14, 13, 13}, {4, 14, 14, 13, 13}},
    {{4, 3, 3, 2, 1}, {4, 3, 3, 13, 13}, {4, 14, 14, 13, 13}, {4, 14, 14, 13, 13}, {4, 14, 14, 13, 13}, {4, 14, 14, 13, 13}},
    {{4, 3, 3, 2, 1}, {4, 3, 3, 15, 16}, {4, 15, 17, 17, 16}, {4, 15, 17, 16, 16}, {4, 15, 17, 16, 16}, {4, 15, 15, 16, 16}},
    {{4, 3, 3, 2, 1}, {4, 3, 0, 16, 16}, {4, 18, 0, 19, 16}, {4, 17, 17, 17, 17}, {4, 17, 17, 17, 17}, {4, 17, 17, 17, 17}},
    {{4, 3, 3, 2, 1}, {4, 3, 18, 19, 19}, {4, 18, 18, 19, 19}, {4, 18, 18, 19, 19}, {4, 18, 18, 19, 19}, {4, 7, 18, 18, 19}},
    {{4, 3, 3, 2, 1}, {4, 3, 2, 19, 19}, {4, 18, 18, 19, 19}, {4, 18, 18, 19, 19}, {4, 7, 18, 19, 19}, {4, 7, 18, 18, 19}}};

// color_value (FuzzyCalc.cpp:625-651): the first trapezoid of the class that
// gives a positive membership, else 0
__device__ __forceinline__ double hue_membership(int cls, double x) {
  double res = 0;
#pragma unroll
  for (int k = 0; k < 11; k++) {
    const HueTrapezoid t = kHueTraps[k];
    if (t.cls != cls) continue;
    if (!res && (x < t.a || x > t.d)) res = 0;
    else if (!res && x >= t.b && x <= t.c) res = 1;
    else if (!res && x < t.b) res = (x - t.a) / (t.b - t.a);
    else if (!res && x > t.c) res = 1 - (x - t.c) / (t.d - t.c);
    if (res > 0) return res;
  }
  return res;
}
// triangular memberships of FuzzyCalc.cpp:654-750: rising while x < m (lt)
// or x <= m (le), breakpoints of FuzzyCalc.h:60-108
__device__ __forceinline__ double tri_lt(double x, double a, double m, double e) {
  if (x < a || x > e) return 0;
  if (x < m) return (x - a) / (m - a);
  return 1 - ((x - m) / (e - m));
}
__device__ __forceinline__ double tri_le(double x, double a, double m, double e) {
  if (x < a || x > e) return 0;
  if (x <= m) return (x - a) / (m - a);
  return 1 - ((x - m) / (e - m));
}

// FindColor for one (h, s, v) byte triple: 1..19
__device__ __forceinline__ int fuzzy_class(int hb, int sb, int vb) {
  const double h = hb, s = sb, v = vb;
  double mh[10];
#pragma unroll
  for (int a = 0; a < 10; a++) mh[a] = hue_membership(kRuleHue[a], h);
  const double ms[6] = {s < 0 ? 0 : tri_lt(s, -0.1, 0, 40), tri_lt(s, 10, 30, 70),    tri_lt(s, 30, 70, 120),
                        tri_lt(s, 60, 120, 180),             tri_lt(s, 120, 180, 240), tri_le(s, 180, 240, 240.1)};
  const double mv[5] = {v < 0 ? 0 : tri_lt(v, -0.1, 0, 60), tri_lt(v, 0, 60, 120), tri_lt(v, 60, 120, 180),
                        tri_lt(v, 120, 180, 240), tri_le(v, 200, 240, 240.1)};
  double sum[kColorClasses + 1];
#pragma unroll
  for (int c = 0; c <= kColorClasses; c++) sum[c] = 0;
  // CalculateRules2 for every class at once: each class's rules in table order
#pragma unroll
  for (int a = 0; a < 10; a++)
#pragma unroll
    for (int k = 0; k < 6; k++)
#pragma unroll
      for (int l = 0; l < 5; l++) {
        const int cls = kRuleClass[a][k][l];
        if (cls == 0) continue;
        double m = mh[a];
        if (m == 0) continue;
        m = m * ms[k];
        if (m == 0) continue;
        m = m * mv[l];
        sum[cls] = sum[cls] + m;
      }
  double best = 0;
  int res = 2;  // light grey when no class has a positive membership
#pragma unroll
  for (int c = 1; c <= kColorClasses; c++)
    if (sum[c] > best) {
      best = sum[c];
      res = c;
    }
  return res;
}

}  // namespace wnd
