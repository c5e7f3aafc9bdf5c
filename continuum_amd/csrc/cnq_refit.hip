// Triangle-hierarchy refit on the GPU, one height at a time: a pass over
// the leaves computes each leaf's box and 27 coverage samples from the new
// pose, then one pass per height computes the internal nodes of that height
// (a leaf has height 0, a node one more than its taller child) from their
// children, which earlier passes finished. Each node is written once; the
// passes are ordered by the stream, so no node waits on another's atomics.
//
// Same results, bit for bit, as the reference's reverse sweep
// (crates/continuum-gpu/src/hierarchy.rs:579-619 `refit_triangle_hierarchy`):
// a leaf's box is the min/max of its triangles' padded boxes (`padded` :762)
// and its samples sum max(area normal . d, 0) over its triangles in item
// order; a parent's box and samples combine its left then right child.
#include "cnq_kernels.h"

namespace cnq {

__device__ __forceinline__ f3 coverage_dir(int n) {
  f3 d{(float)(n / 9) - 1.0f, (float)(n / 3 % 3) - 1.0f, (float)(n % 3) - 1.0f};
  float len = sqrtf(dot3(d, d));
  float rcp = 1.0f / len;
  // glam's normalize_or_zero: the zero vector (n = 13) stays zero.
  return (rcp < 3.4e38f && rcp > 0.0f) ? d * rcp : f3{0.0f, 0.0f, 0.0f};
}

__device__ __forceinline__ float gmx(float a, float b) { return a > b ? a : (b > a ? b : a); }
__device__ __forceinline__ float gmn(float a, float b) { return a < b ? a : (b < a ? b : a); }

__global__ __launch_bounds__(256) void refit_leaf_kernel(RefitArgs a) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_leaves) return;
  const uint32_t n = a.leaves[t];
  const NodeC node = a.blas[n];
  float lo[3] = {3.40282347e38f, 3.40282347e38f, 3.40282347e38f};
  float hi[3] = {-3.40282347e38f, -3.40282347e38f, -3.40282347e38f};
  float cov[27];
#pragma unroll
  for (int k = 0; k < 27; k++) cov[k] = 0.0f;
  // Up to eight triangles at a time: every load of a group is issued before
  // the first store, then the triangles are folded in item order.
  const uint32_t end = node.a + node.b;
  for (uint32_t k0 = node.a; k0 < end; k0 += 8) {
    const uint32_t g = end - k0 < 8u ? end - k0 : 8u;
    uint32_t idx[8];
#pragma unroll
    for (int i = 0; i < 8; i++) idx[i] = (uint32_t)i < g ? a.blas_items[k0 + i] : 0u;
    f3 tv[8][3];
#pragma unroll
    for (int i = 0; i < 8; i++) {
      if ((uint32_t)i < g) {
        const float* tr = a.tris + 12 * (size_t)idx[i];
        tv[i][0] = ld3(tr), tv[i][1] = ld3(tr + 4), tv[i][2] = ld3(tr + 8);
      } else {
        tv[i][0] = tv[i][1] = tv[i][2] = f3{0.0f, 0.0f, 0.0f};
      }
    }
#pragma unroll
    for (int i = 0; i < 8; i++) {
      if ((uint32_t)i >= g) break;
      const uint32_t k = k0 + i;
      const f3 v0 = tv[i][0], v1 = tv[i][1], v2 = tv[i][2];
      float* tl = a.tri_leaf + 9 * (size_t)k;
      tl[0] = v0.x, tl[1] = v0.y, tl[2] = v0.z, tl[3] = v1.x, tl[4] = v1.y, tl[5] = v1.z;
      tl[6] = v2.x, tl[7] = v2.y, tl[8] = v2.z;
      // padded(mesh_bounds([t]))
      float blo[3], bhi[3];
      blo[0] = gmn(gmn(gmn(3.40282347e38f, v0.x), v1.x), v2.x);
      blo[1] = gmn(gmn(gmn(3.40282347e38f, v0.y), v1.y), v2.y);
      blo[2] = gmn(gmn(gmn(3.40282347e38f, v0.z), v1.z), v2.z);
      bhi[0] = gmx(gmx(gmx(-3.40282347e38f, v0.x), v1.x), v2.x);
      bhi[1] = gmx(gmx(gmx(-3.40282347e38f, v0.y), v1.y), v2.y);
      bhi[2] = gmx(gmx(gmx(-3.40282347e38f, v0.z), v1.z), v2.z);
      float m0 = gmx(fabsf(blo[0]), fabsf(bhi[0])), m1 = gmx(fabsf(blo[1]), fabsf(bhi[1])),
            m2 = gmx(fabsf(blo[2]), fabsf(bhi[2]));
      float pad = 1e-5f * gmx(gmx(gmx(m0, m1), m2), 1.0f);
#pragma unroll
      for (int c = 0; c < 3; c++) {
        lo[c] = gmn(lo[c], blo[c] - pad);
        hi[c] = gmx(hi[c], bhi[c] + pad);
      }
      f3 an = cross3(v1 - v0, v2 - v0) * 0.5f;
#pragma unroll
      for (int k2 = 0; k2 < 27; k2++) cov[k2] += gmx(dot3(an, coverage_dir(k2)), 0.0f);
    }
  }
  NodeC* out = a.blas + n;
  out->lo[0] = lo[0], out->lo[1] = lo[1], out->lo[2] = lo[2];
  out->hi[0] = hi[0], out->hi[1] = hi[1], out->hi[2] = hi[2];
  float* oc = a.blas_cov + 27 * (size_t)n;
#pragma unroll
  for (int k = 0; k < 27; k++) oc[k] = cov[k];
}

// Internal node p from its left then right child.
__device__ __forceinline__ void refit_node(const RefitArgs& a, uint32_t p) {
  const uint32_t left = a.blas[p].a;
  const NodeC l = a.blas[left], r = a.blas[left + 1];
  NodeC* out = a.blas + p;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    out->lo[c] = gmn(gmn(3.40282347e38f, l.lo[c]), r.lo[c]);
    out->hi[c] = gmx(gmx(-3.40282347e38f, l.hi[c]), r.hi[c]);
  }
  const float* cl = a.blas_cov + 27 * (size_t)left;
  const float* cr = a.blas_cov + 27 * (size_t)(left + 1);
  float* oc = a.blas_cov + 27 * (size_t)p;
#pragma unroll
  for (int k = 0; k < 27; k++) oc[k] = cl[k] + cr[k];
}

// Internal nodes nodes[0, n), all of one height.
__global__ __launch_bounds__(256) void refit_inner_kernel(RefitArgs a, const uint32_t* __restrict__ nodes, uint32_t n) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) refit_node(a, nodes[t]);
}

// The small heights [h0, levels) in one workgroup, a barrier between
// heights (the workgroup's waves share one CU, so a barrier orders their
// global writes and reads).
__global__ __launch_bounds__(1024) void refit_tail_kernel(RefitArgs a, const uint32_t* __restrict__ inner,
                                                          const uint32_t* __restrict__ level_off, uint32_t h0,
                                                          uint32_t levels) {
  for (uint32_t h = h0; h < levels; h++) {
    const uint32_t b = level_off[h], e = level_off[h + 1];
    for (uint32_t i = b + threadIdx.x; i < e; i += blockDim.x) refit_node(a, inner[i]);
    __syncthreads();
  }
}

hipError_t launch_refit(const RefitArgs& a, const uint32_t* inner, const uint32_t* level_off,
                        const uint32_t* level_off_d, uint32_t n_levels, hipStream_t stream) {
  if (a.n_leaves == 0) return hipSuccess;
  hipLaunchKernelGGL(refit_leaf_kernel, dim3((a.n_leaves + 255) / 256), dim3(256), 0, stream, a);
  hipError_t e = hipGetLastError();
  // Heights from `tail` up hold at most 1,024 nodes each: one workgroup
  // does them all, in one launch instead of one a height.
  uint32_t tail = n_levels;
  while (tail > 0 && level_off[tail] - level_off[tail - 1] <= 1024u) tail--;
  if (n_levels - tail < 2) tail = n_levels;
  for (uint32_t h = 0; h < tail && e == hipSuccess; h++) {
    const uint32_t n = level_off[h + 1] - level_off[h];
    if (n == 0) continue;
    hipLaunchKernelGGL(refit_inner_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, a, inner + level_off[h], n);
    e = hipGetLastError();
  }
  if (tail < n_levels && e == hipSuccess) {
    hipLaunchKernelGGL(refit_tail_kernel, dim3(1), dim3(1024), 0, stream, a, inner, level_off_d, tail, n_levels);
    e = hipGetLastError();
  }
  return e;
}

}  // namespace cnq
