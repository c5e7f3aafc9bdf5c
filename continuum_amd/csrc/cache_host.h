// Host side of the lighting cache (see cache_host.cpp).
#pragma once
#include <stdint.h>

namespace cnq {
namespace host {

constexpr uint32_t BLOCK_PLANES_HOST = 16;

// For each block of 64 texels (8 floats each: point, normal) with bounds
// `boxes` (6 floats: lo, hi), 16 planes (n.x, n.y, n.z, w; inside n.p <= w)
// enclosing every texel's cone toward the rect `light` (a LightGpu, 16
// floats): the hull of the block's box and the light, pushed out by 1e-3.
// Unused planes are zero; a block no texel of which can receive the light
// gets the plane (0, 0, 0, -1), which excludes everything.
void block_planes(const float* texels, const float* boxes, uint32_t n_blocks, const float* light, float* planes);

}  // namespace host
}  // namespace cnq
