// TEST INFRASTRUCTURE ONLY. The oracle is a CPU restatement of the reference's
// cone-visibility path. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the checker: the product library
// (continuum_amd) never links, calls or falls back to anything here.
//
// Scalar f32 arithmetic with the WGSL specification's definitions where the
// spec gives one (normalize = e / length(e), mix = a*(1-t) + b*t,
// clamp = min(max(e, lo), hi), cross by its component formula, dot summed left
// to right). Built with -ffp-contract=off so no multiply-add is fused.
#pragma once
#include <cmath>
#include <cstdint>

namespace orc {

struct V2 {
  float x, y;
};
struct V3 {
  float x, y, z;
};

inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
inline V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
inline V2 operator+(V2 a, V2 b) { return V2{a.x + b.x, a.y + b.y}; }
inline V2 operator-(V2 a, V2 b) { return V2{a.x - b.x, a.y - b.y}; }
inline V2 operator*(V2 a, float s) { return V2{a.x * s, a.y * s}; }

inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float dot2(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
inline V3 cross(V3 a, V3 b) {
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float length(V3 a) { return std::sqrt(dot(a, a)); }
// WGSL: normalize(e) = e / length(e).
inline V3 normalize(V3 a) { return a / length(a); }
// WGSL min/max on one component (no NaNs on this path).
inline float fmin_(float a, float b) { return b < a ? b : a; }
inline float fmax_(float a, float b) { return b > a ? b : a; }
inline float clampf(float e, float lo, float hi) { return fmin_(fmax_(e, lo), hi); }
inline V3 vmin(V3 a, V3 b) { return V3{fmin_(a.x, b.x), fmin_(a.y, b.y), fmin_(a.z, b.z)}; }
inline V3 vmax(V3 a, V3 b) { return V3{fmax_(a.x, b.x), fmax_(a.y, b.y), fmax_(a.z, b.z)}; }
inline V3 vabs(V3 a) { return V3{std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
inline float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }

// --- Elementary functions: one published method, used identically by the
// oracle and the HIP kernels so their results agree bit for bit.
// Arctangent: S. L. Moshier, Cephes Math Library (atanf.c): reduction by
// tan(3pi/8) and tan(pi/8), then a degree-9 odd minimax polynomial.
constexpr float kPi = 3.14159265358979323846f;
constexpr float kPio2 = 1.57079632679489661923f;
constexpr float kPio4 = 0.78539816339744830962f;

inline float atan_c(float xin) {
  float x = xin;
  float sgn = 1.0f;
  if (x < 0.0f) {
    sgn = -1.0f;
    x = -x;
  }
  float y;
  if (x > 2.414213562373095f) {
    y = kPio2;
    x = -1.0f / x;
  } else if (x > 0.4142135623730950f) {
    y = kPio4;
    x = (x - 1.0f) / (x + 1.0f);
  } else {
    y = 0.0f;
  }
  float z = x * x;
  y = y + ((((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z -
            3.33329491539e-1f) *
               z * x +
           x);
  return sgn * y;
}

// atan2 by the same method (Cephes atan2f's quadrant rule).
inline float atan2_c(float y, float x) {
  if (x == 0.0f) {
    if (y > 0.0f) return kPio2;
    if (y < 0.0f) return -kPio2;
    return 0.0f;
  }
  if (y == 0.0f) return x > 0.0f ? 0.0f : kPi;
  float w = 0.0f;
  if (x < 0.0f) w = y < 0.0f ? -kPi : kPi;
  return w + atan_c(y / x);
}

// acos(y) = atan2(sqrt((1 - y)(1 + y)), y).
inline float acos_c(float y) { return atan2_c(std::sqrt((1.0f - y) * (1.0f + y)), y); }

// asin(x) = atan2(x, sqrt((1 - x)(1 + x))) for |x| <= 1.
inline float asin_c(float x) { return atan2_c(x, std::sqrt((1.0f - x) * (1.0f + x))); }

// Cosine and tangent: S. L. Moshier, Cephes Math Library (sinf.c `cosf`,
// tanf.c): the argument reduced by multiples of pi/4 (pi/4 in three parts),
// then minimax polynomials on [-pi/4, pi/4].
constexpr float kFourOverPi = 1.27323954473516f;
constexpr float kDP1 = 0.78515625f;
constexpr float kDP2 = 2.4187564849853515625e-4f;
constexpr float kDP3 = 3.77489497744594108e-8f;

inline float cos_c(float xin) {
  float x = xin < 0.0f ? -xin : xin;
  int j = (int)(kFourOverPi * x);
  float y = (float)j;
  if (j & 1) {
    j += 1;
    y += 1.0f;
  }
  j &= 7;
  int sign = 1;
  if (j > 3) {
    j -= 4;
    sign = -sign;
  }
  if (j > 1) sign = -sign;
  x = ((x - y * kDP1) - y * kDP2) - y * kDP3;
  const float z = x * x;
  if (j == 1 || j == 2)
    y = ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * x + x;
  else
    y = ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z - 0.5f * z + 1.0f;
  return sign < 0 ? -y : y;
}

inline float tan_c(float xin) {
  const float x = xin < 0.0f ? -xin : xin;
  int j = (int)(kFourOverPi * x);
  float y = (float)j;
  if (j & 1) {
    j += 1;
    y += 1.0f;
  }
  const float z = ((x - y * kDP1) - y * kDP2) - y * kDP3;
  const float zz = z * z;
  if (x > 1.0e-4f)
    y = (((((9.38540185543e-3f * zz + 3.11992232697e-3f) * zz + 2.44301354525e-2f) * zz + 5.34112807005e-2f) * zz +
          1.33387994085e-1f) *
             zz +
         3.33331568548e-1f) *
            zz * z +
        z;
  else
    y = z;
  if (j & 2) y = -1.0f / y;
  return xin < 0.0f ? -y : y;
}

}  // namespace orc
