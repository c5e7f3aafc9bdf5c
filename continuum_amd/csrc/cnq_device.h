// Device-side geometry of a mask cone for gfx950: vector math, the cone's
// frame and culling tests, and the exact-area accumulation rasterizer.
//
// Semantics follow the reference's GPU kernel,
// crates/continuum-gpu/src/mask_cones.wgsl (cited as "WGSL :line"), with the
// WGSL specification's definitions for normalize (e / length(e)), mix,
// clamp and cross, dot products summed left to right, and no fused
// multiply-adds (the library is built with -ffp-contract=off). Division and
// square root are IEEE correctly rounded (hipcc's default for f32).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cnq {

constexpr uint32_t NONE = 0xffffffffu;
constexpr uint32_t GRID = 8;
constexpr uint32_t BINS = 64;
constexpr float NEAR = 1e-4f;  // WGSL :211
// WGSL's abstract-precision constant 1.0 - 1e-5, rounded once to f32.
constexpr float ONE_MINUS_1EM5 = (float)(1.0 - 1e-5);
// The reference's traversal stack (WGSL :111); a walk deeper than this is
// counted as an overflow.
constexpr uint32_t REF_STACK = 48;

struct f3 {
  float x, y, z;
};
struct f2 {
  float x, y;
};

__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ f2 operator+(f2 a, f2 b) { return f2{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ f2 operator-(f2 a, f2 b) { return f2{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ f2 operator*(f2 a, float s) { return f2{a.x * s, a.y * s}; }

__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float dot2(f2 a, f2 b) { return a.x * b.x + a.y * b.y; }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
  return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float len3(f3 a) { return sqrtf(dot3(a, a)); }
__device__ __forceinline__ f3 norm3(f3 a) { return a / len3(a); }
__device__ __forceinline__ float mn(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float mx(float a, float b) { return b > a ? b : a; }
__device__ __forceinline__ float clampf(float e, float lo, float hi) { return mn(mx(e, lo), hi); }
__device__ __forceinline__ float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ f3 ld3(const float* p) { return f3{p[0], p[1], p[2]}; }

// Arctangent after S. L. Moshier's Cephes atanf: reduction by tan(3pi/8) and
// tan(pi/8), then a degree-9 odd minimax polynomial (relative error below
// 2e-7); atan2 by Cephes atan2f's quadrant rule.
constexpr float kPi = 3.14159265358979323846f;
constexpr float kPio2 = 1.57079632679489661923f;
constexpr float kPio4 = 0.78539816339744830962f;

__device__ __forceinline__ float atan_c(float xin) {
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
  y = y + ((((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z - 3.33329491539e-1f) * z *
               x +
           x);
  return sgn * y;
}

__device__ __forceinline__ float atan2_c(float y, float x) {
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

__device__ __forceinline__ float acos_c(float y) { return atan2_c(sqrtf((1.0f - y) * (1.0f + y)), y); }

// asin(x) = atan2(x, sqrt((1 - x)(1 + x))) for |x| <= 1.
__device__ __forceinline__ float asin_c(float x) { return atan2_c(x, sqrtf((1.0f - x) * (1.0f + x))); }

// Cosine and tangent: S. L. Moshier, Cephes Math Library (sinf.c `cosf`,
// tanf.c): the argument reduced by multiples of pi/4 (pi/4 in three parts),
// then minimax polynomials on [-pi/4, pi/4].
constexpr float kFourOverPi = 1.27323954473516f;
constexpr float kDP1 = 0.78515625f;
constexpr float kDP2 = 2.4187564849853515625e-4f;
constexpr float kDP3 = 3.77489497744594108e-8f;

__device__ __forceinline__ float cos_c(float xin) {
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

__device__ __forceinline__ float tan_c(float xin) {
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

// Exponential after Cephes expf: x = n ln2 + r (ln2 in two parts), a
// degree-6 polynomial for e^r, then scaling by 2^n.
__device__ __forceinline__ float exp_c(float x) {
  if (x > 88.72283905206835f) return __builtin_inff();
  if (x < -103.278929903431851103f) return 0.0f;
  float z = floorf(1.44269504088896341f * x + 0.5f);
  x = x - z * 0.693359375f;
  x = x - z * -2.12194440e-4f;
  int n = (int)z;
  float xx = x * x;
  float p = ((((((1.9875691500e-4f * x + 1.3981999507e-3f) * x + 8.3334519073e-3f) * x + 4.1665795894e-2f) * x +
               1.6666665459e-1f) *
                  x +
              5.0000001201e-1f) *
                 xx +
             x) +
            1.0f;
  return ldexpf(p, n);
}

// A mask cone (WGSL :125-146). Light, sky and guard cones take their
// forward axis as the far plane's normal; a lattice cone (a moving rect
// light's, in the view cache) takes the light's plane.
struct Cone {
  f3 apex, right, up, forward;
  float w0, w1, w2, w3;  // window u0, v0, u1, v1
  f3 far_normal;
  float far_offset;
  f3 baxis;  // bounding circular cone: unit axis, cos and sin of the half-angle
  float bcos, bsin;
  // The receiver's horizon (WGSL :147): read only by sky walks (`SKY`),
  // whose cones cull geometry wholly below it; light cones have none.
  f3 horizon;
};

__device__ __forceinline__ f3 direction(const Cone& c, float s, float t) {
  return c.forward + c.right * (c.w0 + s * (c.w2 - c.w0)) + c.up * (c.w1 + t * (c.w3 - c.w1));
}

__device__ __forceinline__ f3 to_local(const Cone& c, f3 p) {
  f3 d = p - c.apex;
  return f3{dot3(c.right, d), dot3(c.up, d), dot3(c.forward, d)};
}

// WGSL :159-168
__device__ __forceinline__ bool in_bound(const Cone& c, f3 p, float r) {
  if (c.bcos < -1.0f) return true;
  f3 d = p - c.apex;
  float len = len3(d);
  if (len <= r) return true;
  float s = r / len;
  float limit = c.bcos * sqrtf(1.0f - s * s) - c.bsin * s;
  return dot3(c.baxis, d) >= limit * len;
}

// in_bound for a box's bounding sphere, settled by approximate square roots
// and reciprocals where they prove the exact answer: 1 (true), 0 (false),
// 2 (too close to call: use in_bound). v_sqrt_f32 and v_rcp_f32 are within
// 1 ulp, so `la`, `ra` are within 2.4e-7 of the exact len and r relative
// (the margins below allow 4e-6), and with 1 - s^2 >= 1e-3 the
// approximate limit * len is within 2.5e-5 (|cos| + |sin|) len of the exact
// one (the margin allows 1e-4).
__device__ __forceinline__ int in_bound_quick(const Cone& c, f3 lo, f3 hi) {
  if (c.bcos < -1.0f) return 1;
  const f3 e = hi - lo;
  const f3 d = (lo + hi) * 0.5f - c.apex;
  const float r2 = dot3(e, e), l2 = dot3(d, d);
  if (!(r2 > 1e-30f && l2 > 1e-30f && r2 < 1e30f && l2 < 1e30f)) return 2;
  const float ra = 0.5f * __builtin_amdgcn_sqrtf(r2), la = __builtin_amdgcn_sqrtf(l2);
  if (la <= ra * (1.0f - 4e-6f)) return 1;
  if (!(la > ra * (1.0f + 4e-6f))) return 2;
  const float s = ra * __builtin_amdgcn_rcpf(la);
  const float om = 1.0f - s * s;
  if (!(om >= 1e-3f)) return 2;
  const float rhs = (c.bcos * __builtin_amdgcn_sqrtf(om) - c.bsin * s) * la;
  const float margin = 1e-4f * (fabsf(c.bcos) + fabsf(c.bsin)) * la;
  const float lhs = dot3(c.baxis, d);
  if (lhs >= rhs + margin) return 1;
  if (lhs < rhs - margin) return 0;
  return 2;
}

// WGSL :174-186
__device__ __forceinline__ void bound_cone(Cone& c, f3 axis) {
  float cos_half = 1.0f;
#pragma unroll
  for (uint32_t k = 0; k < 4; k++) {
    f3 d = norm3(direction(c, (float)(k & 1u), (float)(k >> 1u)));
    cos_half = mn(cos_half, dot3(axis, d));
  }
  float cos_b = clampf(cos_half - 1e-5f, -1.0f, 1.0f);
  c.baxis = axis;
  c.bcos = cos_b;
  c.bsin = sqrtf(mx(1.0f - cos_b * cos_b, 0.0f));
}

// WGSL :193-201
__device__ __forceinline__ float lambert_edge_unit(f3 n, f3 an, f3 bn) {
  f3 g = cross3(an, bn);
  float len = len3(g);
  if (len > 1e-12f) return atan2_c(len, dot3(an, bn)) * dot3(n, g / len);
  return 0.0f;
}
__device__ __forceinline__ float lambert_edge(f3 n, f3 a, f3 b) { return lambert_edge_unit(n, norm3(a), norm3(b)); }

__device__ __forceinline__ f3 box_corner(f3 lo, f3 hi, uint32_t k) {
  return f3{(k & 1u) ? hi.x : lo.x, (k & 2u) ? hi.y : lo.y, (k & 4u) ? hi.z : lo.z};
}

// WGSL :234-255; the horizon test only with SKY (light cones have none)
template <bool SKY = false>
__device__ __forceinline__ bool may_touch_box(const Cone& c, f3 lo, f3 hi) {
  // The bounding-cone test first (a cheap bound settles most boxes; both
  // tests must pass, so their order does not change the answer).
  const int quick = in_bound_quick(c, lo, hi);
  if (quick == 0) return false;
  bool far_out = true, near_out = true, out_u0 = true, out_u1 = true, out_v0 = true, out_v1 = true;
  bool below = SKY;
#pragma unroll
  for (uint32_t k = 0; k < 8; k++) {
    f3 corner = box_corner(lo, hi, k);
    if (c.far_offset - dot3(c.far_normal, corner) >= 0.0f) far_out = false;
    if (SKY && dot3(c.horizon, corner - c.apex) >= 0.0f) below = false;
    f3 p = to_local(c, corner);
    if (p.z >= NEAR) near_out = false;
    if (p.x - c.w0 * p.z >= 0.0f) out_u0 = false;
    if (c.w2 * p.z - p.x >= 0.0f) out_u1 = false;
    if (p.y - c.w1 * p.z >= 0.0f) out_v0 = false;
    if (c.w3 * p.z - p.y >= 0.0f) out_v1 = false;
  }
  if (far_out || near_out || out_u0 || out_u1 || out_v0 || out_v1 || below) return false;
  return quick == 1 || in_bound(c, (lo + hi) * 0.5f, 0.5f * len3(hi - lo));
}

// WGSL :213-232 (the horizon test only with SKY)
template <bool SKY = false>
__device__ __forceinline__ bool may_touch(const Cone& c, f3 v0, f3 v1, f3 v2) {
  float f0 = c.far_offset - dot3(c.far_normal, v0);
  float f1 = c.far_offset - dot3(c.far_normal, v1);
  float f2 = c.far_offset - dot3(c.far_normal, v2);
  if (f0 < 0.0f && f1 < 0.0f && f2 < 0.0f) return false;
  if (SKY && dot3(c.horizon, v0 - c.apex) < 0.0f && dot3(c.horizon, v1 - c.apex) < 0.0f &&
      dot3(c.horizon, v2 - c.apex) < 0.0f)
    return false;
  f3 p0 = to_local(c, v0), p1 = to_local(c, v1), p2 = to_local(c, v2);
  if (p0.z < NEAR && p1.z < NEAR && p2.z < NEAR) return false;
  if (p0.x - c.w0 * p0.z < 0.0f && p1.x - c.w0 * p1.z < 0.0f && p2.x - c.w0 * p2.z < 0.0f) return false;
  if (c.w2 * p0.z - p0.x < 0.0f && c.w2 * p1.z - p1.x < 0.0f && c.w2 * p2.z - p2.x < 0.0f) return false;
  if (p0.y - c.w1 * p0.z < 0.0f && p1.y - c.w1 * p1.z < 0.0f && p2.y - c.w1 * p2.z < 0.0f) return false;
  if (c.w3 * p0.z - p0.y < 0.0f && c.w3 * p1.z - p1.y < 0.0f && c.w3 * p2.z - p2.y < 0.0f) return false;
  return true;
}

// ---- Per-query working memory: the 68-cell accumulation buffer in LDS,
// lane-interleaved (cell s of lane l at s * 64 + l), so every access is
// bank-conflict free whatever cell each lane touches, and 70 slots for clip
// polygons (a 3D polygon of up to 10 vertices at slots 0-29, its clip
// target at 30-59; 2D polygons at 30-49 and 50-69 once the 3D clip is done;
// a splat's hull at 30-45) and the horizon-clipped quads of the weights.
constexpr uint32_t ACC_CELLS = 68;
constexpr uint32_t POLY_SLOTS = 70;
// The clipper's workspace lives in the lane's private memory (scratch, cached
// in L1/L2; consecutive slots, stride 1): only triangles that cross the
// window, the near or the far plane use it, and keeping it out of LDS leaves
// 17 KB of LDS per wave, so two waves fit on each SIMD. The cone kernel's
// one-wave variant keeps it in LDS instead, lane-interleaved (stride 64):
// 35 KB a wave, one wave per SIMD, no scratch round trips in the clips.
constexpr uint32_t POLY_STRIDE = 1;
constexpr uint32_t POLY_STRIDE_LDS = 64;
// Waves per SIMD the walk kernels are compiled for (their registers are
// limited to 256 a lane; the LDS, 17 KB a wave, allows it).
constexpr int WALK_WAVES = 2;
constexpr uint32_t P3A = 0, P3B = 30, Q2A = 30, Q2B = 50;
constexpr uint32_t MAX_POLY = 10;  // WGSL :11

struct Lane {
  float* acc;       // this lane's accumulation cell 0
  float* poly;      // this lane's polygon slot 0
  uint32_t pstride;  // polygon slot i at poly[i * pstride]
  uint32_t span_lo, span_hi;
  // The accumulator is in LDS: an LDS atomic add (ds_add_f32) needs no read
  // round trip, and rounds as v_add_f32 does (round to nearest even,
  // denormals kept: checked bit for bit on the device). A lane's adds to its
  // own cells take effect in program order, so each cell sums its terms in
  // the same order as the plain read-add-write.
  __device__ __forceinline__ void add(uint32_t cell, float v) {
    __hip_atomic_fetch_add((__attribute__((address_space(3))) float*)(acc + cell * 64u), v, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __device__ __forceinline__ float get(uint32_t cell) const { return acc[cell * 64u]; }
  __device__ __forceinline__ void set(uint32_t cell, float v) { acc[cell * 64u] = v; }
  __device__ __forceinline__ f3 p3(uint32_t base, uint32_t i) const {
    const float* p = poly + (base + 3u * i) * pstride;
    return f3{p[0], p[pstride], p[2 * pstride]};
  }
  __device__ __forceinline__ void put3(uint32_t base, uint32_t i, f3 v) {
    float* p = poly + (base + 3u * i) * pstride;
    p[0] = v.x;
    p[pstride] = v.y;
    p[2 * pstride] = v.z;
  }
  __device__ __forceinline__ f2 p2(uint32_t base, uint32_t i) const {
    const float* p = poly + (base + 2u * i) * pstride;
    return f2{p[0], p[pstride]};
  }
  __device__ __forceinline__ void put2(uint32_t base, uint32_t i, f2 v) {
    float* p = poly + (base + 2u * i) * pstride;
    p[0] = v.x;
    p[pstride] = v.y;
  }
};

// WGSL :297-348
__device__ __forceinline__ void draw_edge(Lane& acc, f2 q0, f2 q1, float weight) {
  if (fabsf(q0.y - q1.y) <= 1.1920929e-7f) return;
  float dir = 1.0f;
  f2 p0 = q0, p1 = q1;
  if (q0.y > q1.y) {
    dir = -1.0f;
    p0 = q1;
    p1 = q0;
  }
  float dxdy = (p1.x - p0.x) / (p1.y - p0.y);
  float x = p0.x;
  uint32_t yc = (uint32_t)ceilf(p1.y);
  uint32_t y_end = yc < GRID ? yc : GRID;
  const uint32_t y_begin = (uint32_t)p0.y;
  // The rows' span, once per edge (the rows run upward), and the signed
  // weight: dir is +-1, so (dy dir) w and dy (dir w) are the same float.
  if (y_begin < y_end) {
    acc.span_lo = acc.span_lo < y_begin * GRID ? acc.span_lo : y_begin * GRID;
    const uint32_t hi = (y_end - 1u) * GRID + GRID + 1u;
    acc.span_hi = acc.span_hi > hi ? acc.span_hi : hi;
  }
  const float dir_weight = dir * weight;
  for (uint32_t y = y_begin; y < y_end; y++) {
    uint32_t row = y * GRID;
    float dy = mn((float)(y + 1u), p1.y) - mx((float)y, p0.y);
    float x_next = clampf(x + dxdy * dy, 0.0f, (float)GRID);
    float d = dy * dir_weight;
    float x0 = mn(x, x_next);
    float x1 = mx(x, x_next);
    float x0_floor = floorf(x0);
    uint32_t x0i = (uint32_t)x0_floor;
    float x1_ceil = ceilf(x1);
    uint32_t x1i = (uint32_t)x1_ceil;
    if (x1i <= x0i + 1u) {
      float xmf = 0.5f * (x + x_next) - x0_floor;
      acc.add(row + x0i, d - d * xmf);
      acc.add(row + x0i + 1u, d * xmf);
    } else {
      float s = 1.0f / (x1 - x0);
      float x0f = x0 - x0_floor;
      float a0 = 0.5f * s * (1.0f - x0f) * (1.0f - x0f);
      float x1f = x1 - x1_ceil + 1.0f;
      float am = 0.5f * s * x1f * x1f;
      acc.add(row + x0i, d * a0);
      if (x1i == x0i + 2u) {
        acc.add(row + x0i + 1u, d * (1.0f - a0 - am));
      } else {
        float a1 = s * (1.5f - x0f);
        acc.add(row + x0i + 1u, d * (a1 - a0));
        for (uint32_t xi = x0i + 2u; xi < x1i - 1u; xi++) acc.add(row + xi, d * s);
        float a2 = a1 + (float)(x1i - x0i - 3u) * s;
        acc.add(row + x1i - 1u, d * (1.0f - a2 - am));
      }
      acc.add(row + x1i, d * am);
    }
    x = x_next;
  }
}

__device__ __forceinline__ f2 clamp_grid(f2 v) {
  const float g = (float)GRID;
  return f2{clampf(v.x, 0.0f, g), clampf(v.y, 0.0f, g)};
}

// WGSL :264-278: keeps dot(normal, x) <= offset; crossings from the kept end.
// Polygon `src` (n vertices) to `dst`; returns the new count (at most 10).
__device__ __forceinline__ uint32_t clip_plane3(Lane& L, uint32_t src, uint32_t n, f3 normal, float offset,
                                                uint32_t dst) {
  uint32_t m = 0;
  if (n == 0) return 0;
  // The source's vertices in one batch of loads (slots past n re-read vertex 0).
  f3 v[MAX_POLY];
#pragma unroll
  for (uint32_t i = 0; i < MAX_POLY; i++) v[i] = L.p3(src, i < n ? i : 0u);
  const f3 first = v[0];
  f3 a = first;
  float da = offset - dot3(normal, a);
#pragma unroll
  for (uint32_t i = 0; i < MAX_POLY; i++) {
    if (i < n) {
      const f3 b = i + 1 < n ? v[i + 1 < MAX_POLY ? i + 1 : 0] : first;
      const float db = offset - dot3(normal, b);
      if (da >= 0.0f && m < MAX_POLY) L.put3(dst, m++, a);
      if (da >= 0.0f && db < 0.0f && m < MAX_POLY) L.put3(dst, m++, a + (b - a) * (da / (da - db)));
      if (da < 0.0f && db >= 0.0f && m < MAX_POLY) L.put3(dst, m++, b + (a - b) * (db / (db - da)));
      a = b;
      da = db;
    }
  }
  return m;
}

// WGSL :280-294, in the projected plane. The source polygon's vertices are
// read in one batch of loads before the clip (slots past `n` re-read vertex
// 0 and go unused), so a pass waits on memory once, not once per vertex;
// the arithmetic and the order of the outputs are the loop's own.
__device__ __forceinline__ uint32_t clip_line2(Lane& L, uint32_t src, uint32_t n, f2 normal, float offset,
                                               uint32_t dst) {
  uint32_t m = 0;
  if (n == 0) return 0;
  f2 v[MAX_POLY];
#pragma unroll
  for (uint32_t i = 0; i < MAX_POLY; i++) v[i] = L.p2(src, i < n ? i : 0u);
  const f2 first = v[0];
  f2 a = first;
  float da = offset - dot2(normal, a);
#pragma unroll
  for (uint32_t i = 0; i < MAX_POLY; i++) {
    if (i < n) {
      const f2 b = i + 1 < n ? v[i + 1 < MAX_POLY ? i + 1 : 0] : first;
      const float db = offset - dot2(normal, b);
      if (da >= 0.0f && m < MAX_POLY) L.put2(dst, m++, a);
      if (da >= 0.0f && db < 0.0f && m < MAX_POLY) L.put2(dst, m++, a + (b - a) * (da / (da - db)));
      if (da < 0.0f && db >= 0.0f && m < MAX_POLY) L.put2(dst, m++, b + (a - b) * (db / (db - da)));
      a = b;
      da = db;
    }
  }
  return m;
}

// The four clips to [0, 8]^2 (WGSL :432-437: x >= 0, x <= 8, y >= 0, y <= 8,
// in that order), ping-ponging between the two 2D slots; `at` is where the
// polygon is, and is where it ends up.
__device__ __forceinline__ uint32_t clip_to_grid(Lane& L, uint32_t& at, uint32_t n) {
  const float g = (float)GRID;
  const uint32_t other = at == Q2A ? Q2B : Q2A;
#pragma unroll 1
  for (uint32_t k = 0; k < 4; k++) {
    const f2 normal{k == 0 ? -1.0f : (k == 1 ? 1.0f : 0.0f), k == 2 ? -1.0f : (k == 3 ? 1.0f : 0.0f)};
    const bool odd = (k & 1u) != 0;
    n = clip_line2(L, odd ? other : at, n, normal, odd ? g : 0.0f, odd ? at : other);
  }
  return n;
}

// Vertex `k` (uniform) of a polygon held in registers.
__device__ __forceinline__ f2 pick2(const f2 (&v)[MAX_POLY], uint32_t k) {
  f2 r = v[0];
#pragma unroll
  for (uint32_t j = 1; j < MAX_POLY; j++) r = k == j ? v[j] : r;
  return r;
}

// Draws the polygon's edges in order (its vertices read in one batch).
__device__ __forceinline__ void draw_poly(Lane& L, uint32_t at, uint32_t n, float weight) {
  f2 v[MAX_POLY];
#pragma unroll
  for (uint32_t i = 0; i < MAX_POLY; i++) v[i] = L.p2(at, i < n ? i : 0u);
  const f2 first = clamp_grid(v[0]);
  f2 a = first;
  for (uint32_t i = 0; i < n; i++) {
    const f2 b = i + 1 < n ? clamp_grid(pick2(v, i + 1)) : first;
    draw_edge(L, a, b, weight);
    a = b;
  }
}

__device__ __forceinline__ bool in_window(const Cone& c, f3 v) {
  return !(v.x < c.w0 * v.z || v.x > c.w2 * v.z || v.y < c.w1 * v.z || v.y > c.w3 * v.z);
}

__device__ __forceinline__ f2 project(const Cone& c, f3 v, float su, float sv) {
  return f2{(v.x / v.z - c.w0) * su, (v.y / v.z - c.w1) * sv};
}

__device__ __forceinline__ bool in_grid(f2 q) {
  const float g = (float)GRID;
  return !(q.x < 0.0f || q.y < 0.0f || q.x > g || q.y > g);
}

// The general path of `rasterize` (WGSL :358-431) up to the grid clip: the
// triangle as a polygon, clipped by the far plane (`far_clip`) and the near
// plane where it crosses them, then either projected (wholly in the window)
// or the window's square clipped by the planes through the apex and each of
// its edges. Leaves the 2D polygon at `at` and returns its vertex count (0:
// nothing to draw); `grid_clip`: a vertex lies outside the grid. Without a
// far clip the caller passes the triangle's local corners (l0, l1, l2); where
// all three lie in front of the near plane (the common case) the edge planes
// come from them in registers, and the 3D polygon never goes to memory.
__device__ __forceinline__ uint32_t general_polygon(const Cone& c, Lane& L, f3 v0, f3 v1, f3 v2, bool far_clip,
                                                    f3 l0, f3 l1, f3 l2, float su, float sv, uint32_t& at,
                                                    bool& grid_clip) {
  const float g = (float)GRID;
  grid_clip = false;
  at = Q2A;
  uint32_t n = 3;
  bool all_near_ok = true, all_in_window = true, in_regs = false;
  if (!far_clip) {
    all_near_ok = l0.z >= NEAR && l1.z >= NEAR && l2.z >= NEAR;
    all_in_window = in_window(c, l0) && in_window(c, l1) && in_window(c, l2);
    in_regs = all_near_ok;
    if (!all_near_ok) {
      L.put3(P3A, 0, l0);
      L.put3(P3A, 1, l1);
      L.put3(P3A, 2, l2);
    }
  } else {
    L.put3(P3A, 0, v0);
    L.put3(P3A, 1, v1);
    L.put3(P3A, 2, v2);
    n = clip_plane3(L, P3A, n, c.far_normal, c.far_offset, P3B);
    for (uint32_t i = 0; i < n; i++) {
      const f3 v = to_local(c, L.p3(P3B, i));
      L.put3(P3A, i, v);
      if (v.z < NEAR) all_near_ok = false;
      if (!in_window(c, v)) all_in_window = false;
    }
  }
  if (!all_near_ok) {
    n = clip_plane3(L, P3A, n, f3{0.0f, 0.0f, -1.0f}, -NEAR, P3B);
    for (uint32_t i = 0; i < n; i++) L.put3(P3A, i, L.p3(P3B, i));
  }
  if (n < 3) return 0;
  uint32_t qn;
  if (all_near_ok && all_in_window) {
    qn = n;
    for (uint32_t i = 0; i < n; i++) {
      const f3 p = in_regs ? (i == 0 ? l0 : (i == 1 ? l1 : l2)) : L.p3(P3A, i);
      L.put2(Q2A, i, project(c, p, su, sv));
    }
  } else {
    // The window's square clipped by the planes through the apex and each
    // edge, with the whole triangle's winding (WGSL :376-425).
    const f3 t0 = far_clip ? to_local(c, v0) : l0, t1 = far_clip ? to_local(c, v1) : l1,
             t2 = far_clip ? to_local(c, v2) : l2;
    const float d = dot3(t0, cross3(t1, t2));
    if (d == 0.0f) return 0;
    const float side = d > 0.0f ? 1.0f : -1.0f;
    qn = 4;
    L.put2(Q2A, 0, f2{0.0f, 0.0f});
    L.put2(Q2A, 2, f2{g, g});
    if (side > 0.0f) {
      L.put2(Q2A, 1, f2{g, 0.0f});
      L.put2(Q2A, 3, f2{0.0f, g});
    } else {
      L.put2(Q2A, 1, f2{0.0f, g});
      L.put2(Q2A, 3, f2{g, 0.0f});
    }
    for (uint32_t i = 0; i < n; i++) {
      f3 p, q;
      if (in_regs) {
        p = i == 0 ? l0 : (i == 1 ? l1 : l2);
        q = i == 0 ? l1 : (i == 1 ? l2 : l0);
      } else {
        p = L.p3(P3A, i);
        q = L.p3(P3A, i + 1 == n ? 0 : i + 1);
      }
      const f3 e = cross3(p, q);
      const float len = len3(e);
      if (len == 0.0f) continue;
      const f3 k = e * (side / len);
      const uint32_t to = at == Q2A ? Q2B : Q2A;
      qn = clip_line2(L, at, qn, f2{-k.x / su, -k.y / sv}, k.x * c.w0 + k.y * c.w1 + k.z, to);
      at = to;
    }
  }
  f2 v[MAX_POLY];
#pragma unroll
  for (uint32_t i = 0; i < MAX_POLY; i++) v[i] = L.p2(at, i < qn ? i : 0u);
#pragma unroll
  for (uint32_t i = 0; i < MAX_POLY; i++)
    if (i < qn && !in_grid(v[i])) grid_clip = true;
  return qn;
}

// WGSL :350-451. A triangle wholly in front of the near plane, before the far
// plane and inside the window projects to a triangle inside the grid, drawn
// straight from registers; any other case builds its polygon in the lane's
// workspace, and every such polygon goes through one grid clip and one draw.
__device__ __forceinline__ bool rasterize(const Cone& c, Lane& L, f3 v0, f3 v1, f3 v2, float weight) {
  const float f0 = c.far_offset - dot3(c.far_normal, v0);
  const float f1 = c.far_offset - dot3(c.far_normal, v1);
  const float f2_ = c.far_offset - dot3(c.far_normal, v2);
  const bool far_clip = f0 < 0.0f || f1 < 0.0f || f2_ < 0.0f;
  const float su = (float)GRID / (c.w2 - c.w0);
  const float sv = (float)GRID / (c.w3 - c.w1);
  f3 l0{0.0f, 0.0f, 0.0f}, l1 = l0, l2 = l0;
  uint32_t at = Q2A, qn = 0;
  bool grid_clip = false, built = false;
  if (!far_clip) {
    l0 = to_local(c, v0), l1 = to_local(c, v1), l2 = to_local(c, v2);
    if (l0.z >= NEAR && l1.z >= NEAR && l2.z >= NEAR && in_window(c, l0) && in_window(c, l1) && in_window(c, l2)) {
      const f2 q0 = project(c, l0, su, sv), q1 = project(c, l1, su, sv), q2 = project(c, l2, su, sv);
      if (in_grid(q0) && in_grid(q1) && in_grid(q2)) {
        draw_edge(L, q0, q1, weight);
        draw_edge(L, q1, q2, weight);
        draw_edge(L, q2, q0, weight);
        return true;
      }
      // Rounding put a corner a hair outside the grid: clip as the general path does.
      L.put2(Q2A, 0, q0);
      L.put2(Q2A, 1, q1);
      L.put2(Q2A, 2, q2);
      qn = 3;
      grid_clip = true;
      built = true;
    }
  }
  if (!built) qn = general_polygon(c, L, v0, v1, v2, far_clip, l0, l1, l2, su, sv, at, grid_clip);
  if (grid_clip) qn = clip_to_grid(L, at, qn);
  if (qn < 3) return false;
  draw_poly(L, at, qn, weight);
  return true;
}

}  // namespace cnq
