// Host side of the lighting cache: each lightmap block's bounding volume
// toward the light, as up to 16 planes (crates/continuum-gpu/src/lightmap.rs
// :256-333, `Lightmap::block_planes`). Computed once per light, in f32 with
// the reference's vector arithmetic (glam: normalize = v * (1 / |v|), dot
// summed left to right), so the planes are bit-identical to the reference's.
#include "cache_host.h"

#include <cmath>
#include <cstring>

namespace cnq {
namespace host {

namespace {

struct P3 {
  float x, y, z;
};
inline P3 add(P3 a, P3 b) { return P3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline P3 sub(P3 a, P3 b) { return P3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline P3 scale(P3 a, float s) { return P3{a.x * s, a.y * s, a.z * s}; }
inline float dotp(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline P3 crossp(P3 a, P3 b) { return P3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
inline P3 unit(P3 a) { return scale(a, 1.0f / std::sqrt(dotp(a, a))); }
inline P3 at(const float* p) { return P3{p[0], p[1], p[2]}; }
inline bool same(P3 a, P3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

constexpr float MARGIN = 1e-3f;

// Whether any texel of the block can receive direct light from the rect.
bool block_lit(const float* texels, P3 center, P3 corner, P3 emits, bool one_sided) {
  for (int t = 0; t < 64; t++) {
    const P3 p = at(texels + 8 * t), n = at(texels + 8 * t + 4);
    if (n.x == 0.0f && n.y == 0.0f && n.z == 0.0f) continue;
    if (!(dotp(n, sub(center, p)) > -1e-3f)) continue;
    if (one_sided && !(dotp(emits, sub(p, corner)) > -1e-3f)) continue;
    return true;
  }
  return false;
}

}  // namespace

void block_planes(const float* texels, const float* boxes, uint32_t n_blocks, const float* light, float* planes) {
  const P3 corner = at(light), eu = at(light + 4), ev = at(light + 8);
  const bool one_sided = light[3] > 0.5f;
  const P3 center = add(corner, scale(add(eu, ev), 0.5f));
  const P3 emits = unit(crossp(eu, ev));
  const P3 light_pts[4] = {corner, add(corner, eu), add(corner, ev), add(add(corner, eu), ev)};
  struct Plane {
    P3 n;
    float w;
  };
  Plane found[220];  // at most one per triple of 12 points
  for (uint32_t b = 0; b < n_blocks; b++) {
    float* out = planes + 64 * (size_t)b;
    std::memset(out, 0, 64 * sizeof(float));
    uint32_t count = 0;
    if (!block_lit(texels + 512 * (size_t)b, center, corner, emits, one_sided)) {
      out[3] = -1.0f;
      continue;
    }
    // The box's distinct corners, then the light's four.
    const float* bx = boxes + 6 * (size_t)b;
    P3 pts[12];
    uint32_t np = 0;
    for (int k = 0; k < 8; k++) {
      const P3 c{bx[(k & 1) ? 3 : 0], bx[(k & 2) ? 4 : 1], bx[(k & 4) ? 5 : 2]};
      bool dup = false;
      for (uint32_t q = 0; q < np; q++) dup = dup || same(pts[q], c);
      if (!dup) pts[np++] = c;
    }
    for (const P3& c : light_pts) pts[np++] = c;
    // Every plane through three of the points with all of them on one side
    // is a face of their hull; keep the distinct ones, the first 16 found.
    for (uint32_t i = 0; i < np; i++)
      for (uint32_t j = i + 1; j < np; j++)
        for (uint32_t k = j + 1; k < np; k++) {
          P3 n = crossp(sub(pts[j], pts[i]), sub(pts[k], pts[i]));
          if (dotp(n, n) < 1e-14f) continue;
          n = unit(n);
          float d = dotp(n, pts[i]);
          float below = 0.0f, above = 0.0f;
          for (uint32_t q = 0; q < np; q++) {
            const float h = dotp(n, pts[q]) - d;
            below = std::fmin(below, h);
            above = std::fmax(above, h);
          }
          if (!(above <= 1e-6f)) {
            if (!(below >= -1e-6f)) continue;
            n = P3{-n.x, -n.y, -n.z};
            d = -d;
          }
          bool known = false;
          for (uint32_t q = 0; q < count && !known; q++)
            known = dotp(found[q].n, n) > 1.0f - 1e-6f && std::fabs(found[q].w - MARGIN - d) < 1e-6f;
          if (!known) found[count++] = Plane{n, d + MARGIN};
        }
    // Fewer planes enclose more: past the 16th they are dropped.
    for (uint32_t q = 0; q < count && q < BLOCK_PLANES_HOST; q++) {
      out[4 * q] = found[q].n.x;
      out[4 * q + 1] = found[q].n.y;
      out[4 * q + 2] = found[q].n.z;
      out[4 * q + 3] = found[q].w;
    }
  }
}

}  // namespace host
}  // namespace cnq
