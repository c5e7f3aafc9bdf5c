// The C ABI (include/cnq.h): scene upload and validation, query launches,
// refit, diagnostics. Host code; the kernels live in cnq_cone.hip,
// cnq_primary.hip and cnq_refit.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/cnq.h"
#include "bvh_build.h"
#include "cache_host.h"
#include "cnq_internal.h"

using cnq::NodeC;
using cnq::NONE;

// The face values of cnq_sky_irradiance_device_noalloc (6 per receiver, at
// most CNQ_SKY_NOALLOC_MAX receivers a call): a static array in the
// library's code object, one copy per GPU the library runs on, held until
// the process ends. Host code never passes this name to a kernel: it is a
// host-side stand-in, and each device's address comes from
// hipGetSymbolAddress at cnq_device_create (cnq_device::sky_scratch).
static __device__ __align__(256) float scratch[6ull * CNQ_SKY_NOALLOC_MAX];

namespace cnq {
namespace detail {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(CNQ_HIP_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}



__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(fill_u32_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

// Per object: the min/max of its triangles' vertices (`mesh_bounds`,
// hierarchy.rs:751). Block (m, c) reduces part c of object m's vertices into
// part[6 * (m * chunks + c) ..]; mesh_bounds_merge_kernel then reduces the
// parts. Min and max are exact, so the grouping does not change the box
// (only the sign of a zero, which the padding that follows erases).
__global__ __launch_bounds__(256) void mesh_bounds_kernel(const float* __restrict__ tris,
                                                          const uint32_t* __restrict__ ranges, uint32_t chunks,
                                                          float* part) {
  const uint32_t m = blockIdx.x, c = blockIdx.y;
  const uint32_t first = ranges[4 * m], count = ranges[4 * m + 1];
  const uint64_t verts = 3ull * count, per = (verts + chunks - 1) / chunks;
  const uint64_t v0 = per * c, v1 = v0 + per < verts ? v0 + per : verts;
  float lo[3] = {3.40282347e38f, 3.40282347e38f, 3.40282347e38f};
  float hi[3] = {-3.40282347e38f, -3.40282347e38f, -3.40282347e38f};
  for (uint64_t v = v0 + threadIdx.x; v < v1; v += blockDim.x) {
    const float* p = tris + 4 * (3ull * first + v);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      lo[k] = p[k] < lo[k] ? p[k] : lo[k];
      hi[k] = p[k] > hi[k] ? p[k] : hi[k];
    }
  }
  __shared__ float s[6][256];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    s[k][threadIdx.x] = lo[k];
    s[3 + k][threadIdx.x] = hi[k];
  }
  __syncthreads();
  for (uint32_t w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
#pragma unroll
      for (int k = 0; k < 3; k++) {
        float a = s[k][threadIdx.x + w];
        if (a < s[k][threadIdx.x]) s[k][threadIdx.x] = a;
        float b = s[3 + k][threadIdx.x + w];
        if (b > s[3 + k][threadIdx.x]) s[3 + k][threadIdx.x] = b;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 6; k++) part[6 * ((size_t)m * chunks + c) + k] = s[k][0];
}

// Block m: object m's parts (at most 128), one a lane, reduced in LDS.
__global__ __launch_bounds__(128) void mesh_bounds_merge_kernel(const float* __restrict__ part, uint32_t chunks,
                                                                float* out) {
  const uint32_t m = blockIdx.x, c = threadIdx.x;
  __shared__ float s[6][128];
  for (int k = 0; k < 6; k++)
    s[k][c] = c < chunks ? part[6 * ((size_t)m * chunks + c) + k] : (k < 3 ? 3.40282347e38f : -3.40282347e38f);
  __syncthreads();
  for (uint32_t w = 64; w > 0; w >>= 1) {
    if (c < w) {
      for (int k = 0; k < 3; k++) {
        const float a = s[k][c + w], b = s[3 + k][c + w];
        if (a < s[k][c]) s[k][c] = a;
        if (b > s[3 + k][c]) s[3 + k][c] = b;
      }
    }
    __syncthreads();
  }
  if (c == 0)
    for (int k = 0; k < 6; k++) out[6 * (size_t)m + k] = s[k][0];
}

}  // namespace detail
}  // namespace cnq

using namespace cnq::detail;


namespace cnq {
namespace detail {

NodeC compact(const uint32_t* g, bool object_level) {
  NodeC c;
  std::memcpy(c.lo, g, 12);
  std::memcpy(c.hi, g + 4, 12);
  c.a = g[8];
  c.b = g[9];
  if (object_level && g[10] != 0) c.b |= cnq::NODE_NO_SPLAT;
  return c;
}

int check_options(const cnq_options* o) {
  if (!o) return fail(CNQ_INVALID_ARG, "options are null");
  if (o->hierarchy > 2) return fail(CNQ_INVALID_ARG, "hierarchy must be 0, 1 or 2");
  if (!(o->prefilter >= 0.0f)) return fail(CNQ_INVALID_ARG, "prefilter must be >= 0");
  if (!(o->wide_extent > 0.0f)) return fail(CNQ_INVALID_ARG, "wide_extent must be > 0");
  if (o->profile_stage > 4) return fail(CNQ_INVALID_ARG, "profile_stage must be 0..4");
  return CNQ_OK;
}

int validate(const cnq_scene_desc* d) {
  if (!d) return fail(CNQ_INVALID_ARG, "scene description is null");
  if (d->n_objects == 0 || !d->ranges) return fail(CNQ_INVALID_ARG, "a scene needs at least one range");
  if (d->n_tris && !d->tris) return fail(CNQ_INVALID_ARG, "tris is null");
  if (d->n_tris > 0xffffffffull) return fail(CNQ_CAPACITY, "more than 2^32 - 1 triangles");
  for (uint32_t m = 0; m < d->n_objects; m++) {
    const uint32_t* r = d->ranges + 4 * m;
    if ((uint64_t)r[0] + r[1] > d->n_tris) return fail(CNQ_INVALID_ARG, "range " + std::to_string(m) + " exceeds tris");
    if (r[2] != NONE && r[2] >= d->n_blas) return fail(CNQ_INVALID_ARG, "range " + std::to_string(m) + " root out of blas");
    if (r[3] != NONE && r[3] >= d->n_xforms)
      return fail(CNQ_INVALID_ARG, "range " + std::to_string(m) + " transform out of xforms");
  }
  if (d->n_blas && (!d->blas || !d->blas_items)) return fail(CNQ_INVALID_ARG, "blas arrays are null");
  for (uint32_t n = 0; n < d->n_blas; n++) {
    const uint32_t* g = d->blas + 12 * (size_t)n;
    if (g[11] != 0) return fail(CNQ_INVALID_ARG, "blas node " + std::to_string(n) + ": absent levels are not supported");
    if (g[9] > 0) {
      if (g[9] >= cnq::NODE_NO_SPLAT || (uint64_t)g[8] + g[9] > d->n_blas_items)
        return fail(CNQ_INVALID_ARG, "blas leaf " + std::to_string(n) + " exceeds blas_items");
    } else if (g[8] <= n || (uint64_t)g[8] + 1 >= d->n_blas) {
      return fail(CNQ_INVALID_ARG, "blas node " + std::to_string(n) + ": children must follow it and exist");
    }
  }
  for (uint64_t k = 0; k < d->n_blas_items; k++)
    if (d->blas_items[k] >= d->n_tris) return fail(CNQ_INVALID_ARG, "blas item out of tris");
  if (d->n_nodes == 0 || !d->nodes) return fail(CNQ_INVALID_ARG, "the object hierarchy needs a root");
  for (uint32_t n = 0; n < d->n_nodes; n++) {
    const uint32_t* g = d->nodes + 12 * (size_t)n;
    if (g[9] > 0) {
      if (g[9] >= cnq::NODE_NO_SPLAT || (uint64_t)g[8] + g[9] > d->n_items)
        return fail(CNQ_INVALID_ARG, "object leaf " + std::to_string(n) + " exceeds items");
    } else if (g[8] <= n || (uint64_t)g[8] + 1 >= d->n_nodes) {
      return fail(CNQ_INVALID_ARG, "object node " + std::to_string(n) + ": children must follow it and exist");
    }
  }
  for (uint32_t k = 0; k < d->n_items; k++)
    if (d->items[k] >= d->n_objects) return fail(CNQ_INVALID_ARG, "object item out of ranges");
  if (d->n_xforms && !d->xforms) return fail(CNQ_INVALID_ARG, "xforms is null");
  if (d->n_blas_cov && d->n_blas_cov < d->n_blas) return fail(CNQ_INVALID_ARG, "fewer coverage rows than blas nodes");
  return CNQ_OK;
}

uint32_t log2u(uint32_t v) {
  uint32_t l = 0;
  while ((1u << l) < v) l++;
  return l;
}

// A view's work items for a cone tile tile_w x (64 / tile_w), laid out from
// item `item_off`; returns the items it takes.
uint64_t layout_view(cnq::ViewDesc& v, uint32_t tile_w, uint64_t item_off) {
  const uint32_t tw = tile_w, th = 64u / tile_w;
  v.tiles_x = (v.w + tw - 1) / tw;
  const uint64_t tiles = (uint64_t)v.tiles_x * ((v.h + th - 1) / th);
  v.tiles = (uint32_t)tiles;
  v.item_off = (uint32_t)item_off;
  return tiles * v.n_lights;
}

cnq::QueryParams query_params(cnq_scene* s, const cnq_options* o, uint32_t tile_w) {
  cnq::QueryParams p;
  p.n_objects = s->n_objects;
  p.hierarchy = o->hierarchy;
  p.light_cone_cap = o->light_cone_cap;
  p.wide_cone_cap = o->wide_cone_cap;
  p.profile_stage = o->profile_stage;
  p.object_cov = o->object_cov;
  p.prefilter = o->prefilter;
  p.wide_extent = o->wide_extent;
  p.n_views = 1;
  p.n_items = 0;
  p.item_begin = 0;
  p.tile_w_log2 = log2u(tile_w);
  return p;
}

cnq::ItemArgs item_args(cnq_scene* s, const float* receivers, const float* lights, float* vis, uint64_t* full,
                        uint64_t* touched, float* cov) {
  cnq::ItemArgs a;
  a.s = s->args();
  a.receivers = receivers;
  a.lights = lights;
  a.views = nullptr;
  a.out_visible = vis;
  a.out_full = full;
  a.out_touched = touched;
  a.out_cov = cov;
  a.counters = s->counters.p;
  a.work = s->count_work ? s->work.p : nullptr;
  a.work_counter = nullptr;
  a.order = nullptr;
  a.item_cost = nullptr;
  a.query_flags = s->query_flags;
  a.n_flags = s->query_flags ? s->n_query_flags : 0;
  return a;
}

// Longest items first for a view launched again: the launch records each
// item's time and a one-block pass turns them into the next launch's order
// (cnq::launch_item_order). The order buffer is read and rewritten only on
// the stream the scene first ordered on, so a launch never reads it while it
// changes; launches on other streams run in natural order. Results do not
// depend on the order: every item writes only its own queries.
// A launch being captured into a graph runs in natural order and records no
// times: a graph would keep the order buffers' addresses, which a later,
// larger view frees, and a later, smaller view rewrites only in part.
bool capturing(hipStream_t stream) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

hipError_t order_items(ItemOrder& o, const uint32_t*& order, uint32_t*& item_cost, uint64_t items,
                       const uint64_t key[4], hipStream_t stream) {
  if (items == 0 || items > 0x7fffffffull) return hipSuccess;
  if (capturing(stream)) return hipSuccess;
  if (!o.bound) {
    o.bound = true;
    o.stream = stream;
  }
  if (o.stream != stream) return hipSuccess;
  const bool same = o.ready && o.key[0] == key[0] && o.key[1] == key[1] && o.key[2] == key[2] && o.key[3] == key[3];
  if (!same) {
    if (o.cost.n < items) {
      hipError_t e = hipStreamSynchronize(stream);  // no launch still reads the old buffers
      if (e != hipSuccess) return e;
      o.cost.release(), o.order.release(), o.last.release();
      if ((e = o.cost.alloc(items)) != hipSuccess || (e = o.order.alloc(items)) != hipSuccess ||
          (e = o.last.alloc((items + 3) / 4 * 4)) != hipSuccess) {
        o.cost.release(), o.order.release(), o.last.release();
        o.ready = false;
        return e;
      }
    }
    for (int k = 0; k < 4; k++) o.key[k] = key[k];
    o.ready = false;
    // No times yet (a fill kernel, not a memset, so a captured launch replays it).
    hipError_t e = fill_u32(reinterpret_cast<uint32_t*>(o.last.p), 0xffffffffu, (items + 3) / 4, stream);
    if (e != hipSuccess) return e;
  }
  order = o.ready ? o.order.p : nullptr;
  item_cost = o.cost.p;
  return hipSuccess;
}

hipError_t reorder_items(ItemOrder& o, const uint32_t* item_cost, uint64_t items, hipStream_t stream) {
  if (!item_cost) return hipSuccess;
  hipError_t e = cnq::launch_item_order(o.cost.p, o.last.p, o.order.p, (uint32_t)items, stream);
  if (e == hipSuccess) o.ready = true;
  return e;
}

// Launches the items of `a` in direct mode, at most 2^31 - 1 blocks a launch.
hipError_t launch_items_direct(cnq::ItemArgs a, uint64_t n_items, bool count_work, hipStream_t stream,
                               uint32_t waves) {
  const uint64_t per = 0x7fffffffull;
  if (n_items > per) a.order = nullptr, a.item_cost = nullptr;  // an order covers one launch
  for (uint64_t i0 = 0; i0 < n_items; i0 += per) {
    a.p.item_begin = (uint32_t)i0;
    a.p.n_items = (uint32_t)(i0 + per < n_items ? i0 + per : n_items);
    hipError_t e = cnq::launch_cone_items(a, count_work, 0, stream, waves);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

int make_refit_plan(const std::vector<uint32_t>& child, const std::vector<uint32_t>& roots, RefitPlan& plan) {
  // Breadth first from the roots (parents before children), then heights
  // from the back of that order (children before parents).
  const size_t n_nodes = child.size();
  std::vector<uint32_t> order(roots);
  for (size_t i = 0; i < order.size(); i++) {
    const uint32_t c = child[order[i]];
    if (c == NONE) continue;
    if ((uint64_t)c + 1 >= n_nodes || order.size() + 2 > n_nodes)
      return fail(CNQ_INVALID_ARG, "triangle hierarchy: a child outside the nodes, or a cycle");
    order.push_back(c);
    order.push_back(c + 1);
  }
  std::vector<uint8_t> height(n_nodes, 0);
  std::vector<uint32_t> count(1, 0), leaves;
  for (size_t i = order.size(); i-- > 0;) {
    const uint32_t n = order[i], c = child[n];
    if (c == NONE) {
      leaves.push_back(n);
      continue;
    }
    const uint32_t h = 1u + (height[c] > height[c + 1] ? height[c] : height[c + 1]);
    if (h > 255) return fail(CNQ_INVALID_ARG, "triangle hierarchy deeper than 255 levels");
    height[n] = (uint8_t)h;
    if (count.size() < h + 1) count.resize(h + 1, 0);
    count[h]++;
  }
  // Internal nodes by height (a counting sort), lowest first.
  const uint32_t levels = count.size() > 1 ? (uint32_t)count.size() - 1 : 0;
  plan.level_off.assign(levels + 1, 0);
  for (uint32_t h = 1; h <= levels; h++) plan.level_off[h] = plan.level_off[h - 1] + count[h];
  std::vector<uint32_t> next(plan.level_off), inner(levels ? plan.level_off[levels] : 0);
  for (const uint32_t n : order)
    if (child[n] != NONE) inner[next[height[n] - 1]++] = n;
  hipError_t e = plan.leaves.upload(leaves.data(), leaves.size());
  if (e == hipSuccess) e = plan.inner.upload(inner.data(), inner.size());
  if (e == hipSuccess) e = plan.level_off_d.upload(plan.level_off.data(), plan.level_off.size());
  if (e != hipSuccess) return hip_fail(e, "refit plan");
  return CNQ_OK;
}

hipError_t run_refit(cnq_scene* s, const RefitPlan& plan, hipStream_t stream) {
  cnq::RefitArgs a;
  a.tris = s->tris.p;
  a.blas_items = s->blas_items.p;
  a.blas = s->blas.p;
  a.blas_cov = s->blas_cov.p;
  a.tri_leaf = s->tri_leaf.p;
  a.leaves = plan.leaves.p;
  a.n_leaves = (uint32_t)plan.leaves.n;
  const uint32_t levels = plan.level_off.empty() ? 0 : (uint32_t)plan.level_off.size() - 1;
  return cnq::launch_refit(a, plan.inner.p, plan.level_off.data(), plan.level_off_d.p, levels, stream);
}

int check_object_cov(cnq_scene* s, const cnq_options* o) {
  if (o->object_cov != NONE && (uint64_t)o->object_cov + s->n_nodes > s->n_cov_rows)
    return fail(CNQ_INVALID_ARG, "object_cov + n_nodes exceeds the coverage samples");
  return CNQ_OK;
}

}  // namespace detail
}  // namespace cnq

extern "C" {

void cnq_options_default(cnq_options* o) {
  o->hierarchy = 2;
  o->prefilter = 0.0f;
  o->wide_extent = 8.0f;
  o->light_cone_cap = 4096;
  o->wide_cone_cap = 512;
  o->profile_stage = 0;
  o->object_cov = NONE;
  o->nearest_first = 1;
}

const char* cnq_last_error(void) { return g_last_error.c_str(); }

int cnq_device_create(int ordinal, cnq_device** out) {
  if (!out) return fail(CNQ_INVALID_ARG, "out is null");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(CNQ_NO_DEVICE, "no HIP device");
  if (ordinal < 0 || ordinal >= count) return fail(CNQ_NO_DEVICE, "no HIP device " + std::to_string(ordinal));
  HIP_TRY(hipSetDevice(ordinal), "hipSetDevice");
  int cus = 0;
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ordinal), "compute units");
  float* sky_scratch = nullptr;
  HIP_TRY(hipGetSymbolAddress(reinterpret_cast<void**>(&sky_scratch), HIP_SYMBOL(scratch)), "face scratch address");
  if (!sky_scratch) return fail(CNQ_HIP_ERROR, "face scratch address is null");
  cnq_device* d = new (std::nothrow) cnq_device{ordinal, (uint32_t)(cus > 0 ? cus : 1), sky_scratch};
  if (!d) return fail(CNQ_OOM, "out of host memory");
  *out = d;
  return CNQ_OK;
}

void cnq_device_destroy(cnq_device* dev) { delete dev; }

int cnq_flat_build(const float* tris9, const uint32_t* counts, uint32_t n_meshes, cnq_flat** out) {
  if (!out || (n_meshes && (!counts))) return fail(CNQ_INVALID_ARG, "null argument");
  uint64_t total = 0;
  for (uint32_t m = 0; m < n_meshes; m++) total += counts[m];
  if (total && !tris9) return fail(CNQ_INVALID_ARG, "tris is null");
  if (total > 0x3fffffffull) return fail(CNQ_CAPACITY, "too many triangles");
  cnq_flat* f = new (std::nothrow) cnq_flat;
  if (!f) return fail(CNQ_OOM, "out of host memory");
  try {
    cnq::host::flatten(tris9, counts, n_meshes, f->f);
  } catch (const std::bad_alloc&) {
    delete f;
    return fail(CNQ_OOM, "out of host memory");
  }
  *out = f;
  return CNQ_OK;
}

int cnq_flat_sizes(const cnq_flat* flat, uint64_t sizes[6]) {
  if (!flat || !sizes) return fail(CNQ_INVALID_ARG, "null argument");
  sizes[0] = flat->f.tris.size() / 12;
  sizes[1] = flat->f.ranges.size() / 4;
  sizes[2] = flat->f.blas.size();
  sizes[3] = flat->f.blas_items.size();
  sizes[4] = flat->f.nodes.size();
  sizes[5] = flat->f.items.size();
  return CNQ_OK;
}

int cnq_flat_copy(const cnq_flat* flat, float* tris, uint32_t* ranges, uint32_t* blas, uint32_t* blas_items,
                  float* blas_cov, uint32_t* nodes, uint32_t* items) {
  if (!flat) return fail(CNQ_INVALID_ARG, "null argument");
  const cnq::host::Flat& f = flat->f;
  if (tris) std::memcpy(tris, f.tris.data(), f.tris.size() * 4);
  if (ranges) std::memcpy(ranges, f.ranges.data(), f.ranges.size() * 4);
  if (blas) std::memcpy(blas, f.blas.data(), f.blas.size() * 48);
  if (blas_items) std::memcpy(blas_items, f.blas_items.data(), f.blas_items.size() * 4);
  if (blas_cov) std::memcpy(blas_cov, f.blas_cov.data(), f.blas.size() * 27 * 4);
  if (nodes) std::memcpy(nodes, f.nodes.data(), f.nodes.size() * 48);
  if (items) std::memcpy(items, f.items.data(), f.items.size() * 4);
  return CNQ_OK;
}

void cnq_flat_destroy(cnq_flat* flat) { delete flat; }

int cnq_flat_build_instanced(const float* tris9, const uint32_t* counts, uint32_t n_geometries,
                             const cnq_placement* objects, uint64_t n_objects, int object_prefilter, cnq_flat** out) {
  if (!out || (n_geometries && !counts) || (n_objects && !objects)) return fail(CNQ_INVALID_ARG, "null argument");
  uint64_t total = 0;
  for (uint32_t g = 0; g < n_geometries; g++) total += counts[g];
  if (total && !tris9) return fail(CNQ_INVALID_ARG, "tris is null");
  if (total > 0x3fffffffull || n_objects > 0x3fffffffull) return fail(CNQ_CAPACITY, "too many triangles or objects");
  for (uint64_t k = 0; k < n_objects; k++) {
    if (objects[k].geometry >= n_geometries) return fail(CNQ_INVALID_ARG, "object " + std::to_string(k) + ": no such geometry");
    if (!(objects[k].scale > 0.0f)) return fail(CNQ_INVALID_ARG, "object " + std::to_string(k) + ": scale must be > 0");
  }
  static_assert(sizeof(cnq_placement) == sizeof(cnq::host::Placement), "placement layouts agree");
  cnq_flat* f = new (std::nothrow) cnq_flat;
  if (!f) return fail(CNQ_OOM, "out of host memory");
  try {
    cnq::host::flatten_instanced(tris9, counts, n_geometries, reinterpret_cast<const cnq::host::Placement*>(objects),
                                 n_objects, object_prefilter != 0, f->f, f->xforms, f->object_cov);
  } catch (const std::bad_alloc&) {
    delete f;
    return fail(CNQ_OOM, "out of host memory");
  }
  *out = f;
  return CNQ_OK;
}

int cnq_flat_build_instanced_movers(const float* tris9, const uint32_t* counts, uint32_t n_geometries,
                                    const cnq_placement* objects, uint64_t n_objects, const uint32_t* movable,
                                    uint32_t n_movable, int object_prefilter, cnq_flat** out) {
  if (!out || (n_geometries && !counts) || (n_objects && !objects) || (n_movable && !movable))
    return fail(CNQ_INVALID_ARG, "null argument");
  if (n_movable == 0) return cnq_flat_build_instanced(tris9, counts, n_geometries, objects, n_objects, object_prefilter, out);
  std::vector<uint8_t> seen(n_objects, 0);
  for (uint32_t k = 0; k < n_movable; k++) {
    if (movable[k] >= n_objects || seen[movable[k]]) return fail(CNQ_INVALID_ARG, "movable objects must be distinct objects");
    seen[movable[k]] = 1;
  }
  if (n_movable >= n_objects) return fail(CNQ_INVALID_ARG, "a scene with movers needs a static object");
  uint64_t total = 0;
  for (uint32_t g = 0; g < n_geometries; g++) total += counts[g];
  if (total && !tris9) return fail(CNQ_INVALID_ARG, "tris is null");
  if (total > 0x3fffffffull || n_objects > 0x3fffffffull) return fail(CNQ_CAPACITY, "too many triangles or objects");
  for (uint64_t k = 0; k < n_objects; k++) {
    if (objects[k].geometry >= n_geometries) return fail(CNQ_INVALID_ARG, "object " + std::to_string(k) + ": no such geometry");
    if (!(objects[k].scale > 0.0f)) return fail(CNQ_INVALID_ARG, "object " + std::to_string(k) + ": scale must be > 0");
  }
  cnq_flat* f = new (std::nothrow) cnq_flat;
  if (!f) return fail(CNQ_OOM, "out of host memory");
  try {
    cnq::host::flatten_instanced(tris9, counts, n_geometries, reinterpret_cast<const cnq::host::Placement*>(objects),
                                 n_objects, object_prefilter != 0, f->f, f->xforms, f->object_cov, movable, n_movable,
                                 &f->movers);
  } catch (const std::bad_alloc&) {
    delete f;
    return fail(CNQ_OOM, "out of host memory");
  }
  f->has_movers = true;
  *out = f;
  return CNQ_OK;
}

int cnq_flat_mover_tree(const cnq_flat* flat, cnq_mover_tree* out) {
  if (!flat || !out) return fail(CNQ_INVALID_ARG, "null argument");
  if (!flat->has_movers) return fail(CNQ_INVALID_ARG, "the flat has no movers");
  std::memcpy(out->fixed_root, &flat->movers.fixed_root, 48);
  out->first_node = flat->movers.first_node;
  out->nodes = flat->movers.nodes;
  out->first_item = flat->movers.first_item;
  out->n_objects = (uint32_t)flat->movers.objects.size();
  out->objects = flat->movers.objects.data();
  return CNQ_OK;
}

int cnq_flat_instanced_sizes(const cnq_flat* flat, uint64_t sizes[3]) {
  if (!flat || !sizes) return fail(CNQ_INVALID_ARG, "null argument");
  sizes[0] = flat->xforms.size() / 16;
  sizes[1] = flat->f.blas_cov.size() / 27;
  sizes[2] = flat->object_cov;
  return CNQ_OK;
}

int cnq_flat_copy_instanced(const cnq_flat* flat, float* xforms, float* blas_cov) {
  if (!flat) return fail(CNQ_INVALID_ARG, "null argument");
  if (xforms) std::memcpy(xforms, flat->xforms.data(), flat->xforms.size() * 4);
  if (blas_cov) std::memcpy(blas_cov, flat->f.blas_cov.data(), flat->f.blas_cov.size() * 4);
  return CNQ_OK;
}

int cnq_scene_create(cnq_device* dev, const cnq_scene_desc* d, cnq_scene** out) {
  if (!dev || !out) return fail(CNQ_INVALID_ARG, "null argument");
  int rc = validate(d);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(dev->ordinal), "hipSetDevice");
  cnq_scene* s = new (std::nothrow) cnq_scene;
  if (!s) return fail(CNQ_OOM, "out of host memory");
  s->dev = dev;
  s->n_tris = d->n_tris;
  s->n_objects = d->n_objects;
  s->n_blas = d->n_blas;
  s->n_blas_items = d->n_blas_items;
  for (uint32_t n = 0; n < d->n_blas; n++) {
    const uint32_t b = d->blas[12 * (size_t)n + 9];
    if (b > s->max_leaf) s->max_leaf = b;
  }
  s->n_nodes = d->n_nodes;
  s->n_items = d->n_items;
  s->n_xforms = d->n_xforms;
  s->h_ranges.assign(d->ranges, d->ranges + 4 * (size_t)d->n_objects);
  for (uint32_t m = 0; m < d->n_objects; m++)
    if (d->ranges[4 * m + 3] != NONE) s->instanced = true;
  std::vector<NodeC> blas(d->n_blas), nodes(d->n_nodes);
  std::vector<uint8_t> is_child(d->n_blas, 0);
  s->h_child.assign(d->n_blas, NONE);
  for (uint32_t n = 0; n < d->n_blas; n++) {
    const uint32_t* g = d->blas + 12 * (size_t)n;
    blas[n] = compact(g, false);
    if (g[9] == 0) {
      s->h_child[n] = g[8];
      is_child[g[8]] = is_child[g[8] + 1] = 1;
    }
  }
  std::vector<uint32_t> roots;
  for (uint32_t n = 0; n < d->n_blas; n++)
    if (!is_child[n]) roots.push_back(n);
  for (uint32_t n = 0; n < d->n_nodes; n++) nodes[n] = compact(d->nodes + 12 * (size_t)n, true);
  std::vector<float> tri_leaf(9 * d->n_blas_items);
  for (uint64_t k = 0; k < d->n_blas_items; k++) {
    const float* t = d->tris + 12 * (size_t)d->blas_items[k];
    float* o = tri_leaf.data() + 9 * k;
    o[0] = t[0], o[1] = t[1], o[2] = t[2], o[3] = t[4], o[4] = t[5], o[5] = t[6], o[6] = t[8], o[7] = t[9],
    o[8] = t[10];
  }
  s->n_cov_rows = d->n_blas_cov ? d->n_blas_cov : d->n_blas;
  std::vector<float> zero_cov;
  const float* cov = d->blas_cov;
  if (!cov) {
    zero_cov.assign(27 * (size_t)s->n_cov_rows, 0.0f);
    cov = zero_cov.data();
  }
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = s->tris.upload(d->tris, 12 * d->n_tris);
  if (e == hipSuccess) e = s->tri_leaf.upload(tri_leaf.data(), tri_leaf.size());
  if (e == hipSuccess) e = s->blas_items.upload(d->blas_items, d->n_blas_items);
  if (e == hipSuccess) e = s->ranges.upload(d->ranges, 4 * (size_t)d->n_objects);
  if (e == hipSuccess) e = s->blas.upload(blas.data(), blas.size());
  if (e == hipSuccess) e = s->blas_cov.upload(cov, 27 * (size_t)s->n_cov_rows);
  if (e == hipSuccess) e = s->nodes.upload(nodes.data(), nodes.size());
  if (e == hipSuccess) e = s->items.upload(d->items, d->n_items);
  if (e == hipSuccess) e = s->xforms.upload(d->xforms, 16 * (size_t)d->n_xforms);
  // Mesh bounds in parts of about 8,192 vertices of the largest object (at
  // most 128 parts an object), then the objects' boxes after the parts.
  uint64_t max_count = 0;
  for (uint32_t m = 0; m < d->n_objects; m++) max_count = std::max<uint64_t>(max_count, d->ranges[4 * (size_t)m + 1]);
  s->bound_chunks = (uint32_t)std::min<uint64_t>(128, std::max<uint64_t>(1, (3 * max_count + 8191) / 8192));
  if (e == hipSuccess) e = s->mesh_box.alloc(6 * (size_t)d->n_objects * (s->bound_chunks + 1));
  if (e == hipSuccess) e = s->counters.alloc(4);
  if (e == hipSuccess) e = s->work.alloc(6);
  if (e == hipSuccess) e = hipMemset(s->counters.p, 0, s->counters.bytes());
  if (e == hipSuccess) e = hipMemset(s->work.p, 0, s->work.bytes());
  if (e != hipSuccess) {
    s->release();
    delete s;
    return hip_fail(e, e == hipErrorOutOfMemory ? "scene upload (out of device memory)" : "scene upload");
  }
  if ((rc = make_refit_plan(s->h_child, roots, s->refit))) {
    s->release();
    delete s;
    return rc;
  }
  *out = s;
  return CNQ_OK;
}

void cnq_scene_destroy(cnq_scene* s) {
  if (!s) return;
  (void)hipSetDevice(s->dev->ordinal);
  s->release();
  delete s;
}

uint64_t cnq_scene_bytes(const cnq_scene* s) { return s ? s->bytes() : 0; }

int cnq_irradiance_device(cnq_scene* s, const float* receivers, uint64_t n, const float* lights, uint32_t n_lights,
                          const cnq_options* o, float* vis, uint64_t* full, uint64_t* touched, float* cov,
                          void* stream) {
  if (!s) return fail(CNQ_INVALID_ARG, "scene is null");
  int rc = check_options(o);
  if (rc) return rc;
  if ((rc = check_object_cov(s, o))) return rc;
  if (n == 0 || n_lights == 0) return CNQ_OK;
  if (!receivers || !lights || !vis || !full || !touched) return fail(CNQ_INVALID_ARG, "null buffer");
  if (n > 0xffffffffull || (n + 63) / 64 * n_lights > 0xffffffffull)
    return fail(CNQ_CAPACITY, "too many queries for one call");
  // The receivers as an n x 1 view, tiles of 64 consecutive receivers.
  cnq::ItemArgs a = item_args(s, receivers, lights, vis, full, touched, cov);
  a.p = query_params(s, o, 64);
  cnq::ViewDesc& v = a.v0;
  v.recv_off = 0;
  v.out_off = 0;
  v.w = (uint32_t)n;
  v.h = 1;
  v.light_off = 0;
  v.n_lights = n_lights;
  v.ptile_off = 0;
  const uint64_t items = layout_view(v, 64, 0);
  HIP_TRY(launch_items_direct(a, items, s->count_work, (hipStream_t)stream, 2), "cone query launch");
  return CNQ_OK;
}

int cnq_view_irradiance_device(cnq_scene* s, const float* receivers, uint32_t width, uint32_t height,
                               const float* lights, uint32_t n_lights, const cnq_options* o, uint32_t tile_w,
                               float* vis, uint64_t* full, uint64_t* touched, void* stream) {
  if (!s) return fail(CNQ_INVALID_ARG, "scene is null");
  int rc = check_options(o);
  if (rc) return rc;
  if ((rc = check_object_cov(s, o))) return rc;
  if (tile_w == 0 || tile_w > 64 || (tile_w & (tile_w - 1))) return fail(CNQ_INVALID_ARG, "tile_w must be 1..64, a power of two");
  if ((uint64_t)width * height == 0 || n_lights == 0) return CNQ_OK;
  if (!receivers || !lights || !vis || !full || !touched) return fail(CNQ_INVALID_ARG, "null buffer");
  cnq::ItemArgs a = item_args(s, receivers, lights, vis, full, touched, nullptr);
  a.p = query_params(s, o, tile_w);
  cnq::ViewDesc& v = a.v0;
  v.recv_off = 0;
  v.out_off = 0;
  v.w = width;
  v.h = height;
  v.light_off = 0;
  v.n_lights = n_lights;
  v.ptile_off = 0;
  const uint64_t items = layout_view(v, tile_w, 0);
  if (items > 0xffffffffull) return fail(CNQ_CAPACITY, "too many queries for one call");
  const uint64_t key[4] = {((uint64_t)width << 32) | height, ((uint64_t)n_lights << 32) | tile_w, items,
                           (uint64_t)(uintptr_t)receivers};
  const uint32_t* order = nullptr;
  uint32_t* cost = nullptr;
  HIP_TRY(order_items(s->order, order, cost, items, key, (hipStream_t)stream), "item order buffers");
  a.order = order;
  a.item_cost = cost;
  HIP_TRY(launch_items_direct(a, items, s->count_work, (hipStream_t)stream, 2), "cone query launch");
  HIP_TRY(reorder_items(s->order, a.item_cost, items, (hipStream_t)stream), "item order");
  return CNQ_OK;
}

int cnq_irradiance(cnq_scene* s, const float* receivers, uint64_t n, const float* lights, uint32_t n_lights,
                   const cnq_options* o, float* out_visible, uint64_t* out_full, uint64_t* out_touched,
                   float* out_cov) {
  if (!s) return fail(CNQ_INVALID_ARG, "scene is null");
  int rc = check_options(o);
  if (rc) return rc;
  if (n == 0 || n_lights == 0) return CNQ_OK;
  if (!receivers || !lights || !out_visible || !out_full || !out_touched) return fail(CNQ_INVALID_ARG, "null buffer");
  HIP_TRY(hipSetDevice(s->dev->ordinal), "hipSetDevice");
  const uint64_t q = n * (uint64_t)n_lights;
  DevBuf<float> r, l, v, c;
  DevBuf<uint64_t> f, t;
  auto cleanup = [&]() { r.release(), l.release(), v.release(), c.release(), f.release(), t.release(); };
  hipError_t e = r.upload(receivers, 8 * n);
  if (e == hipSuccess) e = l.upload(lights, 16 * (size_t)n_lights);
  if (e == hipSuccess) e = v.alloc(q);
  if (e == hipSuccess) e = f.alloc(q);
  if (e == hipSuccess) e = t.alloc(q);
  if (e == hipSuccess && out_cov) e = c.alloc(64 * q);
  if (e != hipSuccess) {
    cleanup();
    return hip_fail(e, "query buffers");
  }
  rc = cnq_irradiance_device(s, r.p, n, l.p, n_lights, o, v.p, f.p, t.p, out_cov ? c.p : nullptr, nullptr);
  if (rc == CNQ_OK) {
    e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(out_visible, v.p, 4 * q, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(out_full, f.p, 8 * q, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(out_touched, t.p, 8 * q, hipMemcpyDeviceToHost);
    if (e == hipSuccess && out_cov) e = hipMemcpy(out_cov, c.p, 256 * q, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = hip_fail(e, "query results");
  }
  cleanup();
  return rc;
}

void cnq_sky_options_default(cnq_sky_options* sky) {
  if (!sky) return;
  sky->max_combine = 0;
  sky->horizon_cull = 1;
  sky->prefilter_scale = 1.0f;
}

}  // extern "C"

namespace {

// The sky launch over n receivers: a list (w 0) or a w x h view in tiles of
// tile_w x 64 / tile_w pixels.
int sky_device(cnq_scene* s, const float* receivers, uint64_t n, uint32_t w, uint32_t h, uint32_t tile_w,
               const cnq_options* o, const cnq_sky_options* sky, float* out, float* out_faces, uint64_t* out_full,
               void* stream) {
  if (!s) return fail(CNQ_INVALID_ARG, "scene is null");
  int rc = check_options(o);
  if (rc) return rc;
  if ((rc = check_object_cov(s, o))) return rc;
  if (n == 0) return CNQ_OK;
  if (!receivers || !out) return fail(CNQ_INVALID_ARG, "null buffer");
  cnq::SkyArgs a;
  a.n = (uint32_t)n;
  a.w = w;
  a.h = w ? h : 1u;
  a.tile_w_log2 = w ? log2u(tile_w) : 6u;
  if (n > 0xffffffffull || cnq::sky_items(a) > 0x7fffffffull)
    return fail(CNQ_CAPACITY, "too many receivers for one call");
  cnq_sky_options so;
  cnq_sky_options_default(&so);
  if (sky) so = *sky;
  HIP_TRY(hipSetDevice(s->dev->ordinal), "hipSetDevice");
  if (!out_faces) {
    if (capturing((hipStream_t)stream))
      return fail(CNQ_INVALID_ARG, "a captured sky launch needs its own out_faces");
    if (!s->sky_faces.p || s->sky_faces.n < 6 * n) {
      HIP_TRY(hipDeviceSynchronize(), "wait before the face buffer grows");  // no launch still reads it
      s->sky_faces.release();
      HIP_TRY(s->sky_faces.alloc(6 * n), "sky face buffer");
    }
    out_faces = s->sky_faces.p;
  }
  a.s = s->args();
  a.receivers = receivers;
  a.out = out;
  a.out_faces = out_faces;
  a.out_full = out_full;
  a.work = s->work.p;
  a.p = query_params(s, o, 64);
  a.sky = cnq::SkyParams{so.max_combine, so.horizon_cull, so.prefilter_scale};
  a.order = nullptr;
  a.item_cost = nullptr;
  const uint64_t items = cnq::sky_items(a);
  const uint64_t key[4] = {n, (uint64_t)(uintptr_t)receivers, items, ((uint64_t)w << 32) | (w ? tile_w : 0u)};
  const uint32_t* order = nullptr;
  uint32_t* cost = nullptr;
  HIP_TRY(order_items(s->sky_order, order, cost, items, key, (hipStream_t)stream), "sky order buffers");
  a.order = order;
  a.item_cost = cost;
  HIP_TRY(cnq::launch_sky(a, s->count_work, (hipStream_t)stream), "sky launch");
  HIP_TRY(reorder_items(s->sky_order, a.item_cost, items, (hipStream_t)stream), "sky order");
  return CNQ_OK;
}

}  // namespace

extern "C" {

int cnq_sky_irradiance_device(cnq_scene* s, const float* receivers, uint64_t n, const cnq_options* o,
                              const cnq_sky_options* sky, float* out, float* out_faces, uint64_t* out_full,
                              void* stream) {
  return sky_device(s, receivers, n, 0, 1, 64, o, sky, out, out_faces, out_full, stream);
}

int cnq_sky_view_irradiance_device(cnq_scene* s, const float* receivers, uint32_t width, uint32_t height,
                                   uint32_t tile_w, const cnq_options* o, const cnq_sky_options* sky, float* out,
                                   float* out_faces, uint64_t* out_full, void* stream) {
  if (tile_w == 0 || tile_w > 64 || (tile_w & (tile_w - 1)))
    return fail(CNQ_INVALID_ARG, "tile_w must be 1..64, a power of two");
  return sky_device(s, receivers, (uint64_t)width * height, width, height, tile_w, o, sky, out, out_faces, out_full,
                    stream);
}

int cnq_sky_irradiance_device_noalloc(cnq_scene* s, const float* receivers, uint64_t n, const cnq_options* o,
                                      const cnq_sky_options* sky, float* out, uint64_t* out_full, void* stream) {
  if (!s) return fail(CNQ_INVALID_ARG, "scene is null");
  int rc = check_options(o);
  if (rc) return rc;
  if ((rc = check_object_cov(s, o))) return rc;
  if (n == 0) return CNQ_OK;
  if (!receivers || !out) return fail(CNQ_INVALID_ARG, "null buffer");
  if (n > CNQ_SKY_NOALLOC_MAX)
    return fail(CNQ_CAPACITY, "at most " + std::to_string(CNQ_SKY_NOALLOC_MAX) + " receivers a call without allocation");
  if (!s->dev->sky_scratch) return fail(CNQ_HIP_ERROR, "no face scratch on this device");
  cnq_sky_options so;
  cnq_sky_options_default(&so);
  if (sky) so = *sky;
  HIP_TRY(hipSetDevice(s->dev->ordinal), "hipSetDevice");
  cnq::SkyArgs a;
  a.s = s->args();
  a.receivers = receivers;
  a.out = out;
  a.out_faces = s->dev->sky_scratch;  // every face of every receiver is written before the sum reads it
  a.out_full = out_full;
  a.work = s->work.p;
  a.n = (uint32_t)n;
  a.p = query_params(s, o, 64);
  a.sky = cnq::SkyParams{so.max_combine, so.horizon_cull, so.prefilter_scale};
  a.order = nullptr;  // natural order: no order buffers
  a.item_cost = nullptr;
  HIP_TRY(cnq::launch_sky(a, s->count_work, (hipStream_t)stream), "sky launch");
  return CNQ_OK;
}

int cnq_sky_irradiance(cnq_scene* s, const float* receivers, uint64_t n, const cnq_options* o,
                       const cnq_sky_options* sky, float* out, float* out_faces, uint64_t* out_full) {
  if (!s) return fail(CNQ_INVALID_ARG, "scene is null");
  int rc = check_options(o);
  if (rc) return rc;
  if (n == 0) return CNQ_OK;
  if (!receivers || !out) return fail(CNQ_INVALID_ARG, "null buffer");
  HIP_TRY(hipSetDevice(s->dev->ordinal), "hipSetDevice");
  DevBuf<float> r, v, fa;
  DevBuf<uint64_t> fu;
  auto cleanup = [&]() { r.release(), v.release(), fa.release(), fu.release(); };
  hipError_t e = r.upload(receivers, 8 * n);
  if (e == hipSuccess) e = v.alloc(n);
  if (e == hipSuccess) e = fa.alloc(6 * n);
  if (e == hipSuccess && out_full) e = fu.alloc(6 * n);
  if (e != hipSuccess) {
    cleanup();
    return hip_fail(e, "sky buffers");
  }
  rc = cnq_sky_irradiance_device(s, r.p, n, o, sky, v.p, fa.p, out_full ? fu.p : nullptr, nullptr);
  if (rc == CNQ_OK) {
    e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(out, v.p, 4 * n, hipMemcpyDeviceToHost);
    if (e == hipSuccess && out_faces) e = hipMemcpy(out_faces, fa.p, 24 * n, hipMemcpyDeviceToHost);
    if (e == hipSuccess && out_full) e = hipMemcpy(out_full, fu.p, 48 * n, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = hip_fail(e, "sky results");
  }
  cleanup();
  return rc;
}

int cnq_primary_device(cnq_scene* s, const float* camera, uint32_t width, uint32_t height, const float* lights,
                       uint32_t n_lights, const cnq_options* o, uint32_t* hits, float* receivers, void* stream) {
  if (!s) return fail(CNQ_INVALID_ARG, "scene is null");
  int rc = check_options(o);
  if (rc) return rc;
  if (width == 0 || height == 0) return CNQ_OK;
  if (!camera || !hits || !receivers || (n_lights && !lights)) return fail(CNQ_INVALID_ARG, "null buffer");
  cnq::PrimaryArgs a;
  a.s = s->args();
  a.cameras = camera;
  a.lights = lights;
  a.views = nullptr;
  cnq::ViewDesc& v = a.v0;
  v = cnq::ViewDesc{};
  v.w = width;
  v.h = height;
  v.light_off = 0;
  v.n_lights = n_lights;
  v.ptile_off = 0;
  const uint64_t tiles = (uint64_t)((width + 7) / 8) * ((height + 7) / 8);
  if (tiles > 0x7fffffffull) return fail(CNQ_CAPACITY, "view too large");
  a.n_views = 1;
  a.n_tiles = (uint32_t)tiles;
  a.hierarchy = o->hierarchy;
  a.nearest_first = o->nearest_first;
  a.hits = hits;
  a.receivers = receivers;
  a.counters = s->counters.p + 3;
  a.normals = s->normals.p;
  HIP_TRY(cnq::launch_primary(a, (hipStream_t)stream), "primary launch");
  return CNQ_OK;
}

int cnq_render_view(cnq_scene* s, const float* camera, uint32_t width, uint32_t height, const float* lights,
                    uint32_t n_lights, const cnq_options* o, uint32_t* out_hits, float* out_visible,
                    uint64_t* out_full, uint64_t* out_touched) {
  if (!s) return fail(CNQ_INVALID_ARG, "scene is null");
  int rc = check_options(o);
  if (rc) return rc;
  if (!camera || (n_lights && !lights)) return fail(CNQ_INVALID_ARG, "null buffer");
  HIP_TRY(hipSetDevice(s->dev->ordinal), "hipSetDevice");
  const uint64_t px = (uint64_t)width * height, q = px * n_lights;
  if (px == 0) return CNQ_OK;
  DevBuf<float> cam, l, recv, v;
  DevBuf<uint32_t> h;
  DevBuf<uint64_t> f, t;
  auto cleanup = [&]() { cam.release(), l.release(), recv.release(), v.release(), h.release(), f.release(), t.release(); };
  hipError_t e = cam.upload(camera, 20);
  if (e == hipSuccess) e = l.upload(lights, 16 * (size_t)n_lights);
  if (e == hipSuccess) e = recv.alloc(8 * px);
  if (e == hipSuccess) e = h.alloc(4 * px);
  if (e == hipSuccess) e = v.alloc(q);
  if (e == hipSuccess) e = f.alloc(q);
  if (e == hipSuccess) e = t.alloc(q);
  if (e != hipSuccess) {
    cleanup();
    return hip_fail(e, "view buffers");
  }
  rc = cnq_primary_device(s, cam.p, width, height, l.p, n_lights, o, h.p, recv.p, nullptr);
  if (rc == CNQ_OK && n_lights)
    rc = cnq_view_irradiance_device(s, recv.p, width, height, l.p, n_lights, o, 16, v.p, f.p, t.p, nullptr);
  if (rc == CNQ_OK) {
    e = hipDeviceSynchronize();
    if (e == hipSuccess && out_hits) e = hipMemcpy(out_hits, h.p, 16 * px, hipMemcpyDeviceToHost);
    if (e == hipSuccess && q && out_visible) e = hipMemcpy(out_visible, v.p, 4 * q, hipMemcpyDeviceToHost);
    if (e == hipSuccess && q && out_full) e = hipMemcpy(out_full, f.p, 8 * q, hipMemcpyDeviceToHost);
    if (e == hipSuccess && q && out_touched) e = hipMemcpy(out_touched, t.p, 8 * q, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = hip_fail(e, "view results");
  }
  cleanup();
  return rc;
}

int cnq_scene_set_movers(cnq_scene* s, const cnq_mover_tree* t, const float* geometry_bounds, uint32_t n_geometries,
                         const cnq_placement* placements) {
  if (!s || !t || !geometry_bounds || (t->n_objects && (!t->objects || !placements)))
    return fail(CNQ_INVALID_ARG, "null argument");
  if (!s->instanced) return fail(CNQ_INVALID_ARG, "movers need an instanced scene");
  if (t->n_objects == 0) return fail(CNQ_INVALID_ARG, "no movers");
  if ((uint64_t)t->first_node + t->nodes > s->n_nodes || (uint64_t)t->first_item + t->n_objects > s->n_items ||
      s->n_nodes < 3)
    return fail(CNQ_INVALID_ARG, "mover tree outside the object hierarchy");
  if (t->first_node < 3) return fail(CNQ_INVALID_ARG, "the movers' nodes follow nodes 0-2");
  std::vector<uint32_t> mover_of(s->n_objects, NONE);
  for (uint32_t k = 0; k < t->n_objects; k++) {
    const uint32_t m = t->objects[k];
    if (m >= s->n_objects || mover_of[m] != NONE) return fail(CNQ_INVALID_ARG, "movers must be distinct objects");
    if (s->h_ranges[4 * m + 3] == NONE) return fail(CNQ_INVALID_ARG, "a mover needs a transform");
    if (placements[k].geometry >= n_geometries || !(placements[k].scale > 0.0f))
      return fail(CNQ_INVALID_ARG, "mover " + std::to_string(k) + ": bad placement");
    mover_of[m] = k;
  }
  s->movers = cnq::host::MoverTree{};
  std::memcpy(&s->movers.fixed_root, t->fixed_root, 48);
  s->movers.first_node = t->first_node;
  s->movers.nodes = t->nodes;
  s->movers.first_item = t->first_item;
  s->movers.objects.assign(t->objects, t->objects + t->n_objects);
  s->geometry_bounds.resize(n_geometries);
  for (uint32_t g = 0; g < n_geometries; g++) {
    std::memcpy(s->geometry_bounds[g].lo, geometry_bounds + 6 * (size_t)g, 12);
    std::memcpy(s->geometry_bounds[g].hi, geometry_bounds + 6 * (size_t)g + 3, 12);
  }
  s->mover_placements.assign(reinterpret_cast<const cnq::host::Placement*>(placements),
                             reinterpret_cast<const cnq::host::Placement*>(placements) + t->n_objects);
  s->mover_dirty.assign(t->n_objects, 0);
  s->mover_of = std::move(mover_of);
  s->mover_bounds.resize(t->n_objects);
  for (uint32_t k = 0; k < t->n_objects; k++)
    s->mover_bounds[k] = cnq::host::world_bounds(s->mover_placements[k], s->geometry_bounds[s->mover_placements[k].geometry]);
  s->last_moves.clear();
  s->has_movers = true;
  return CNQ_OK;
}

int cnq_scene_set_transform(cnq_scene* s, uint32_t object, const float* rotation, const float* translation) {
  if (!s || !rotation || !translation) return fail(CNQ_INVALID_ARG, "null argument");
  if (!s->has_movers) return fail(CNQ_INVALID_ARG, "the scene has no movers");
  if (object >= s->n_objects || s->mover_of[object] == NONE)
    return fail(CNQ_INVALID_ARG, "object " + std::to_string(object) + " is not movable");
  cnq::host::Placement& p = s->mover_placements[s->mover_of[object]];
  std::memcpy(p.rotation, rotation, 36);
  std::memcpy(p.translation, translation, 12);
  s->mover_dirty[s->mover_of[object]] = 1;
  return CNQ_OK;
}

int cnq_scene_apply_moves(cnq_scene* s, void* stream) {
  if (!s) return fail(CNQ_INVALID_ARG, "scene is null");
  if (!s->has_movers) return fail(CNQ_INVALID_ARG, "the scene has no movers");
  bool any = false;
  for (uint8_t d : s->mover_dirty) any = any || d;
  s->last_moves.clear();
  if (!any) return CNQ_OK;
  HIP_TRY(hipSetDevice(s->dev->ordinal), "hipSetDevice");
  hipStream_t st = (hipStream_t)stream;
  const cnq::host::MoverTree& t = s->movers;
  std::vector<cnq::host::Box> bounds(t.objects.size());
  for (size_t k = 0; k < t.objects.size(); k++) {
    const cnq::host::Placement& p = s->mover_placements[k];
    bounds[k] = cnq::host::world_bounds(p, s->geometry_bounds[p.geometry]);
  }
  for (size_t k = 0; k < t.objects.size(); k++) {
    if (!s->mover_dirty[k]) continue;
    s->mover_dirty[k] = 0;
    // The move, for the view cache's stale passes (instanced.rs:1284-1306).
    cnq_moved mv;
    mv.object = t.objects[k];
    std::memcpy(mv.old_lo, s->mover_bounds[k].lo, 12);
    std::memcpy(mv.old_hi, s->mover_bounds[k].hi, 12);
    std::memcpy(mv.new_lo, bounds[k].lo, 12);
    std::memcpy(mv.new_hi, bounds[k].hi, 12);
    s->last_moves.push_back(mv);
    s->mover_bounds[k] = bounds[k];
    const cnq::host::Placement& p = s->mover_placements[k];
    const float* r = p.rotation;
    const float rows[16] = {r[0], r[1], r[2], p.translation[0], r[3], r[4], r[5], p.translation[1],
                            r[6], r[7], r[8], p.translation[2], p.scale, 1.0f / p.scale, p.clear, 0.0f};
    const uint32_t x = s->h_ranges[4 * t.objects[k] + 3];
    HIP_TRY(hipMemcpyAsync(s->xforms.p + 16 * (size_t)x, rows, 64, hipMemcpyHostToDevice, st), "mover transform");
    HIP_TRY(hipStreamSynchronize(st), "mover transform");
  }
  cnq::host::NodeGpu root;
  std::vector<cnq::host::NodeGpu> region;
  std::vector<uint32_t> items;
  cnq::host::mover_subtree(t, bounds, root, region, items);
  if (region.size() != t.nodes) return fail(CNQ_INVALID_ARG, "movers' subtree changed shape");
  const cnq::host::NodeGpu top = cnq::host::union_root(t.fixed_root, root);
  const NodeC ctop = compact(reinterpret_cast<const uint32_t*>(&top), true);
  const NodeC croot = compact(reinterpret_cast<const uint32_t*>(&root), true);
  std::vector<NodeC> cregion(region.size());
  for (size_t n = 0; n < region.size(); n++) cregion[n] = compact(reinterpret_cast<const uint32_t*>(&region[n]), true);
  HIP_TRY(hipMemcpyAsync(s->nodes.p, &ctop, sizeof(NodeC), hipMemcpyHostToDevice, st), "root node");
  HIP_TRY(hipMemcpyAsync(s->nodes.p + 2, &croot, sizeof(NodeC), hipMemcpyHostToDevice, st), "movers' root");
  if (!cregion.empty())
    HIP_TRY(hipMemcpyAsync(s->nodes.p + t.first_node, cregion.data(), cregion.size() * sizeof(NodeC),
                           hipMemcpyHostToDevice, st), "movers' nodes");
  HIP_TRY(hipMemcpyAsync(s->items.p + t.first_item, items.data(), items.size() * 4, hipMemcpyHostToDevice, st),
          "movers' items");
  HIP_TRY(hipStreamSynchronize(st), "apply moves");
  return CNQ_OK;
}

int cnq_scene_set_pose(cnq_scene* s, uint32_t geometry, const float* tris9, uint32_t n_tris, void* stream) {
  if (!s || !tris9) return fail(CNQ_INVALID_ARG, "null argument");
  if (!s->has_movers) return fail(CNQ_INVALID_ARG, "poses need a scene with movers (cnq_scene_set_movers)");
  if (geometry >= s->geometry_bounds.size()) return fail(CNQ_INVALID_ARG, "geometry out of range");
  // The geometry's triangles and hierarchy, from one of its movers' ranges.
  uint32_t first = NONE, count = 0, root = NONE;
  for (size_t k = 0; k < s->mover_placements.size(); k++) {
    if (s->mover_placements[k].geometry != geometry) continue;
    const uint32_t* r = &s->h_ranges[4 * (size_t)s->movers.objects[k]];
    first = r[0], count = r[1], root = r[2];
    break;
  }
  if (first == NONE) return fail(CNQ_INVALID_ARG, "no movable object places this geometry");
  if (n_tris != count) return fail(CNQ_INVALID_ARG, "a pose keeps the geometry's triangle count");
  // instanced.rs: deforming geometries are placed by movable objects only.
  for (uint32_t m = 0; m < s->n_objects; m++) {
    const uint32_t* r = &s->h_ranges[4 * (size_t)m];
    if (r[0] == first && r[1] == count && r[2] == root && s->mover_of[m] == NONE)
      return fail(CNQ_INVALID_ARG, "object " + std::to_string(m) + " places the posed geometry but is not movable");
  }
  HIP_TRY(hipSetDevice(s->dev->ordinal), "hipSetDevice");
  hipStream_t st = (hipStream_t)stream;
  // The refit plan of the geometry's hierarchy, made once.
  auto it = s->pose_plans.find(root);
  if (it == s->pose_plans.end() && root != NONE) {
    RefitPlan plan;
    const int rc = make_refit_plan(s->h_child, std::vector<uint32_t>{root}, plan);
    if (rc) {
      plan.release();
      return rc;
    }
    it = s->pose_plans.emplace(root, plan).first;
  }
  // Triangles: xyz of each vertex into the 16-byte rows (w keeps its code).
  if (count)
    HIP_TRY(hipMemcpy2DAsync(s->tris.p + 12 * (size_t)first, 16, tris9, 12, 12, 3 * (size_t)count,
                             hipMemcpyHostToDevice, st), "pose triangles");
  if (root != NONE) HIP_TRY(run_refit(s, it->second, st), "pose refit");
  // Its new bounds (hierarchy.rs:751 mesh_bounds) and its instances moved.
  cnq::host::Box b;
  for (int c = 0; c < 3; c++) b.lo[c] = 3.40282347e38f, b.hi[c] = -3.40282347e38f;
  for (uint64_t v = 0; v < 3 * (uint64_t)count; v++)
    for (int c = 0; c < 3; c++) {
      const float x = tris9[3 * v + c];
      b.lo[c] = x < b.lo[c] ? x : b.lo[c];
      b.hi[c] = x > b.hi[c] ? x : b.hi[c];
    }
  s->geometry_bounds[geometry] = b;
  for (size_t k = 0; k < s->mover_placements.size(); k++)
    if (s->mover_placements[k].geometry == geometry) s->mover_dirty[k] = 1;
  HIP_TRY(hipStreamSynchronize(st), "pose");
  return CNQ_OK;
}

int cnq_scene_last_moves(cnq_scene* s, cnq_moved* out, uint32_t cap, uint32_t* n) {
  if (!s || !n || (cap && !out)) return fail(CNQ_INVALID_ARG, "null argument");
  *n = (uint32_t)s->last_moves.size();
  if (*n > cap) return fail(CNQ_CAPACITY, "more moves than `cap`");
  if (*n) std::memcpy(out, s->last_moves.data(), *n * sizeof(cnq_moved));
  return CNQ_OK;
}

int cnq_scene_refit_device(cnq_scene* s, const float* tris, void* stream) {
  if (!s || !tris) return fail(CNQ_INVALID_ARG, "null argument");
  if (s->instanced) return fail(CNQ_INVALID_ARG, "refit of instanced scenes goes through their geometries");
  hipStream_t st = (hipStream_t)stream;
  if (s->n_tris)
    HIP_TRY(hipMemcpyAsync(s->tris.p, tris, 48 * s->n_tris, hipMemcpyDeviceToDevice, st), "refit copy");
  HIP_TRY(run_refit(s, s->refit, st), "refit launch");
  // The object hierarchy over the new mesh bounds (live.rs:888-892 rebuilds
  // `mesh_hierarchy`): bounds on the device, the small tree on the host.
  float* boxes_d = s->mesh_box.p + 6 * (size_t)s->n_objects * s->bound_chunks;
  hipLaunchKernelGGL(mesh_bounds_kernel, dim3(s->n_objects, s->bound_chunks), dim3(256), 0, st, s->tris.p,
                     s->ranges.p, s->bound_chunks, s->mesh_box.p);
  HIP_TRY(hipGetLastError(), "mesh bounds launch");
  hipLaunchKernelGGL(mesh_bounds_merge_kernel, dim3(s->n_objects), dim3(128), 0, st, s->mesh_box.p, s->bound_chunks,
                     boxes_d);
  HIP_TRY(hipGetLastError(), "mesh bounds merge launch");
  std::vector<float> mb(6 * (size_t)s->n_objects);
  HIP_TRY(hipMemcpyAsync(mb.data(), boxes_d, mb.size() * 4, hipMemcpyDeviceToHost, st), "mesh bounds");
  HIP_TRY(hipStreamSynchronize(st), "refit sync");
  std::vector<cnq::host::Box> boxes(s->n_objects);
  for (uint32_t m = 0; m < s->n_objects; m++) {
    cnq::host::Box b;
    std::memcpy(b.lo, &mb[6 * m], 12);
    std::memcpy(b.hi, &mb[6 * m + 3], 12);
    boxes[m] = cnq::host::padded(b);
  }
  std::vector<cnq::host::NodeGpu> nodes;
  std::vector<uint32_t> items;
  cnq::host::object_hierarchy(boxes, nodes, items);
  if (nodes.size() != s->n_nodes || items.size() != s->n_items)
    return fail(CNQ_INVALID_ARG, "refit: the object hierarchy was not built by mesh_hierarchy");
  std::vector<NodeC> cn(nodes.size());
  for (size_t n = 0; n < nodes.size(); n++) cn[n] = compact(reinterpret_cast<const uint32_t*>(&nodes[n]), true);
  HIP_TRY(hipMemcpyAsync(s->nodes.p, cn.data(), cn.size() * sizeof(NodeC), hipMemcpyHostToDevice, st), "nodes");
  HIP_TRY(hipMemcpyAsync(s->items.p, items.data(), items.size() * 4, hipMemcpyHostToDevice, st), "items");
  HIP_TRY(hipStreamSynchronize(st), "refit sync");
  return CNQ_OK;
}

int cnq_scene_refit(cnq_scene* s, const float* tris) {
  if (!s || !tris) return fail(CNQ_INVALID_ARG, "null argument");
  HIP_TRY(hipSetDevice(s->dev->ordinal), "hipSetDevice");
  DevBuf<float> t;
  hipError_t e = t.upload(tris, 12 * s->n_tris);
  if (e != hipSuccess) {
    t.release();
    return hip_fail(e, "refit upload");
  }
  int rc = cnq_scene_refit_device(s, t.p, nullptr);
  (void)hipDeviceSynchronize();
  t.release();
  return rc;
}

int cnq_scene_read_blas(cnq_scene* s, uint32_t* blas, float* blas_cov) {
  if (!s) return fail(CNQ_INVALID_ARG, "scene is null");
  HIP_TRY(hipSetDevice(s->dev->ordinal), "hipSetDevice");
  HIP_TRY(hipDeviceSynchronize(), "sync");
  if (blas) {
    std::vector<NodeC> c(s->n_blas);
    HIP_TRY(hipMemcpy(c.data(), s->blas.p, c.size() * sizeof(NodeC), hipMemcpyDeviceToHost), "read blas");
    for (uint32_t n = 0; n < s->n_blas; n++) {
      uint32_t* g = blas + 12 * (size_t)n;
      std::memcpy(g, c[n].lo, 12);
      g[3] = 0;
      std::memcpy(g + 4, c[n].hi, 12);
      g[7] = 0;
      g[8] = c[n].a;
      g[9] = c[n].b;
      g[10] = 0;
      g[11] = 0;
    }
  }
  if (blas_cov)
    HIP_TRY(hipMemcpy(blas_cov, s->blas_cov.p, 27 * 4 * (size_t)s->n_blas, hipMemcpyDeviceToHost), "read cov");
  return CNQ_OK;
}

int cnq_scene_counters(cnq_scene* s, int reset, uint32_t counters[3], uint64_t work[6]) {
  if (!s) return fail(CNQ_INVALID_ARG, "scene is null");
  HIP_TRY(hipSetDevice(s->dev->ordinal), "hipSetDevice");
  HIP_TRY(hipDeviceSynchronize(), "sync");
  uint32_t c[4];
  uint64_t w[6];
  HIP_TRY(hipMemcpy(c, s->counters.p, sizeof(c), hipMemcpyDeviceToHost), "read counters");
  HIP_TRY(hipMemcpy(w, s->work.p, sizeof(w), hipMemcpyDeviceToHost), "read work");
  if (counters) {
    counters[0] = c[0];
    counters[1] = c[1];
    counters[2] = c[2] + c[3];
  }
  if (work) std::memcpy(work, w, sizeof(w));
  if (reset) {
    HIP_TRY(hipMemset(s->counters.p, 0, s->counters.bytes()), "reset counters");
    HIP_TRY(hipMemset(s->work.p, 0, s->work.bytes()), "reset work");
  }
  return CNQ_OK;
}

int cnq_scene_set_count_work(cnq_scene* s, int on) {
  if (!s) return fail(CNQ_INVALID_ARG, "scene is null");
  s->count_work = on != 0;
  return CNQ_OK;
}

int cnq_scene_set_normals(cnq_scene* s, const float* normals, uint64_t n_tris) {
  if (!s) return fail(CNQ_INVALID_ARG, "scene is null");
  HIP_TRY(hipSetDevice(s->dev->ordinal), "hipSetDevice");
  if (!normals) {
    HIP_TRY(hipDeviceSynchronize(), "wait before dropping the normals");  // no camera pass still reads them
    s->normals.release();
    return CNQ_OK;
  }
  if (n_tris != s->n_tris) return fail(CNQ_INVALID_ARG, "normals: one row of 9 per triangle of the scene");
  if (!s->normals.p) {
    HIP_TRY(s->normals.alloc(9 * n_tris), "normals buffer");
  } else {
    HIP_TRY(hipDeviceSynchronize(), "wait before replacing the normals");
  }
  HIP_TRY(hipMemcpy(s->normals.p, normals, 36 * n_tris, hipMemcpyHostToDevice), "normals upload");
  return CNQ_OK;
}

int cnq_scene_set_query_flags(cnq_scene* s, uint8_t* flags, uint64_t n) {
  if (!s) return fail(CNQ_INVALID_ARG, "scene is null");
  if (!flags && n) return fail(CNQ_INVALID_ARG, "null flags with a size");
  s->query_flags = n ? flags : nullptr;
  s->n_query_flags = n ? n : 0;
  return CNQ_OK;
}

}  // extern "C"

// ---- Batches: many views in one camera pass and one cone launch.
struct cnq_batch {
  cnq_scene* scene = nullptr;
  uint32_t n_views = 0, n_lights = 0, n_ptiles = 0;
  uint64_t px_total = 0, q_total = 0, n_items = 0;
  std::vector<cnq::ViewDesc> h_views;
  uint32_t layout_tile_w = 0;
  DevBuf<cnq::ViewDesc> views;
  DevBuf<float> cameras, lights, receivers, visible;
  DevBuf<uint32_t> hits, work_counter;
  // The cone pass's items in the order of their last times (longest first):
  // identity after a layout, then rewritten after every cone pass, so a
  // replayed batch starts its heaviest items first.
  DevBuf<uint32_t> item_cost, item_order;
  DevBuf<uint8_t> item_last;
  DevBuf<uint64_t> full, touched;
  hipStream_t own_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  cnq_tuning graph_tuning{};

  void release() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (own_stream) (void)hipStreamDestroy(own_stream);
    exec = nullptr, graph = nullptr, own_stream = nullptr;
    views.release(), cameras.release(), lights.release(), receivers.release(), visible.release();
    hits.release(), work_counter.release(), full.release(), touched.release();
    item_cost.release(), item_order.release(), item_last.release();
  }
};

namespace {

int check_tuning(const cnq_tuning* t) {
  if (!t) return fail(CNQ_INVALID_ARG, "tuning is null");
  if (t->tile_w == 0 || t->tile_w > 64 || (t->tile_w & (t->tile_w - 1)))
    return fail(CNQ_INVALID_ARG, "tile_w must be 1..64, a power of two");
  if (t->waves_per_simd > 2) return fail(CNQ_INVALID_ARG, "waves_per_simd must be 0, 1 or 2");
  if (t->ray_stack != 0 && t->ray_stack != 32 && t->ray_stack != 64)
    return fail(CNQ_INVALID_ARG, "ray_stack must be 0, 32 or 64");
  return CNQ_OK;
}

// Lays the batch's work items out for a cone tile width and uploads the
// view table (synchronously: never during a capture). A captured graph holds
// the view table's and the item buffers' addresses and the old layout, so a
// new layout first waits for the device (no launch still reads them) and
// drops the graph: a replay then fails until the batch is captured again.
int batch_layout(cnq_batch* b, uint32_t tile_w) {
  if (b->layout_tile_w == tile_w) return CNQ_OK;
  std::vector<cnq::ViewDesc> laid(b->h_views);
  uint64_t item = 0;
  for (auto& v : laid) {
    item += layout_view(v, tile_w, item);
    if (item > 0xffffffffull) return fail(CNQ_CAPACITY, "a batch holds at most 2^32 - 1 work items");
  }
  HIP_TRY(hipDeviceSynchronize(), "wait before a new layout");
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->layout_tile_w = 0;
  b->h_views.swap(laid);
  b->n_items = item;
  HIP_TRY(hipMemcpy(b->views.p, b->h_views.data(), b->h_views.size() * sizeof(cnq::ViewDesc), hipMemcpyHostToDevice),
          "view table");
  // A fresh identity order for the new items.
  b->item_cost.release(), b->item_order.release(), b->item_last.release();
  if (b->n_items <= 0x7fffffffull) {
    std::vector<uint32_t> ident(b->n_items);
    for (uint64_t i = 0; i < b->n_items; i++) ident[i] = (uint32_t)i;
    const std::vector<uint8_t> none(b->n_items, 0xffu);
    hipError_t e = b->item_cost.alloc(b->n_items);
    if (e == hipSuccess) e = b->item_order.upload(ident.data(), ident.size());
    if (e == hipSuccess) e = b->item_last.upload(none.data(), none.size());
    if (e != hipSuccess) {
      b->item_cost.release(), b->item_order.release(), b->item_last.release();
      return hip_fail(e, "item order");
    }
  }
  b->layout_tile_w = tile_w;
  return CNQ_OK;
}

// Enqueues the camera pass and the cone pass of a laid-out batch.
int batch_enqueue(cnq_batch* b, const cnq_options* o, const cnq_tuning* t, hipStream_t st) {
  cnq_scene* s = b->scene;
  cnq::PrimaryArgs pa;
  pa.s = s->args();
  pa.cameras = b->cameras.p;
  pa.lights = b->lights.p;
  pa.views = b->views.p;
  pa.v0 = b->h_views[0];
  pa.n_views = b->n_views;
  pa.n_tiles = b->n_ptiles;
  pa.hierarchy = o->hierarchy;
  pa.nearest_first = o->nearest_first;
  pa.hits = b->hits.p;
  pa.receivers = b->receivers.p;
  pa.counters = s->counters.p + 3;
  pa.normals = s->normals.p;
  HIP_TRY(cnq::launch_primary(pa, st, t->ray_stack ? t->ray_stack : 64), "batch camera pass");
  cnq::ItemArgs a = item_args(s, b->receivers.p, b->lights.p, b->visible.p, b->full.p, b->touched.p, nullptr);
  a.p = query_params(s, o, t->tile_w);
  a.p.n_views = b->n_views;
  a.views = b->views.p;
  a.v0 = b->h_views[0];
  a.work_counter = b->work_counter.p;
  a.order = b->item_order.p;
  a.item_cost = b->item_cost.p;
  if (t->persistent_blocks) {
    HIP_TRY(fill_u32(b->work_counter.p, 0, 1, st), "work counter");
    a.p.item_begin = 0;
    a.p.n_items = (uint32_t)b->n_items;
    HIP_TRY(cnq::launch_cone_items(a, s->count_work, t->persistent_blocks, st, t->waves_per_simd), "batch cone pass");
  } else {
    HIP_TRY(launch_items_direct(a, b->n_items, s->count_work, st, t->waves_per_simd), "batch cone pass");
  }
  if (a.item_cost)
    HIP_TRY(cnq::launch_item_order(a.item_cost, b->item_last.p, b->item_order.p, (uint32_t)b->n_items, st),
            "item order");
  return CNQ_OK;
}

}  // namespace

extern "C" {

void cnq_tuning_default(cnq_tuning* t) {
  t->tile_w = 16;
  t->persistent_blocks = 0;
  t->waves_per_simd = 2;
  t->ray_stack = 64;
}

int cnq_batch_create(cnq_scene* s, const cnq_view* views, uint32_t n_views, const float* lights, uint32_t n_lights,
                     cnq_batch** out) {
  if (!s || !views || !out || n_views == 0) return fail(CNQ_INVALID_ARG, "null argument or no views");
  if (n_lights && !lights) return fail(CNQ_INVALID_ARG, "lights is null");
  HIP_TRY(hipSetDevice(s->dev->ordinal), "hipSetDevice");
  cnq_batch* b = new (std::nothrow) cnq_batch;
  if (!b) return fail(CNQ_OOM, "out of host memory");
  b->scene = s;
  b->n_views = n_views;
  b->n_lights = n_lights;
  b->h_views.resize(n_views);
  uint64_t px = 0, q = 0, ptiles = 0;
  std::vector<float> cams(20 * (size_t)n_views);
  for (uint32_t i = 0; i < n_views; i++) {
    const cnq_view& cv = views[i];
    if (cv.width == 0 || cv.height == 0 || (uint64_t)cv.light_first + cv.light_count > n_lights) {
      delete b;
      return fail(CNQ_INVALID_ARG, "view " + std::to_string(i) + ": empty, or lights outside the batch's");
    }
    cnq::ViewDesc& v = b->h_views[i];
    v = cnq::ViewDesc{};
    v.recv_off = px;
    v.out_off = q;
    v.w = cv.width;
    v.h = cv.height;
    v.light_off = cv.light_first;
    v.n_lights = cv.light_count;
    v.ptile_off = (uint32_t)ptiles;
    px += (uint64_t)cv.width * cv.height;
    q += (uint64_t)cv.width * cv.height * cv.light_count;
    ptiles += (uint64_t)((cv.width + 7) / 8) * ((cv.height + 7) / 8);
    std::memcpy(&cams[20 * (size_t)i], cv.camera, 80);
  }
  if (ptiles > 0x7fffffffull) {
    delete b;
    return fail(CNQ_CAPACITY, "too many pixels in one batch");
  }
  b->px_total = px;
  b->q_total = q;
  b->n_ptiles = (uint32_t)ptiles;
  hipError_t e = b->views.alloc(n_views);
  if (e == hipSuccess) e = b->cameras.upload(cams.data(), cams.size());
  if (e == hipSuccess) e = b->lights.upload(lights, 16 * (size_t)n_lights);
  if (e == hipSuccess) e = b->receivers.alloc(8 * px);
  if (e == hipSuccess) e = b->hits.alloc(4 * px);
  if (e == hipSuccess) e = b->visible.alloc(q);
  if (e == hipSuccess) e = b->full.alloc(q);
  if (e == hipSuccess) e = b->touched.alloc(q);
  if (e == hipSuccess) e = b->work_counter.alloc(1);
  if (e != hipSuccess) {
    b->release();
    delete b;
    return hip_fail(e, "batch buffers");
  }
  int rc = batch_layout(b, 8);
  if (rc) {
    b->release();
    delete b;
    return rc;
  }
  *out = b;
  return CNQ_OK;
}

void cnq_batch_destroy(cnq_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->scene->dev->ordinal);
  b->release();
  delete b;
}

uint64_t cnq_batch_queries(const cnq_batch* b) { return b ? b->q_total : 0; }

int cnq_batch_run(cnq_batch* b, const cnq_options* o, const cnq_tuning* t, void* stream) {
  if (!b) return fail(CNQ_INVALID_ARG, "batch is null");
  int rc = check_options(o);
  if (rc || (rc = check_tuning(t)) || (rc = check_object_cov(b->scene, o))) return rc;
  if ((rc = batch_layout(b, t->tile_w))) return rc;
  return batch_enqueue(b, o, t, (hipStream_t)stream);
}

int cnq_batch_capture(cnq_batch* b, const cnq_options* o, const cnq_tuning* t) {
  if (!b) return fail(CNQ_INVALID_ARG, "batch is null");
  int rc = check_options(o);
  if (rc || (rc = check_tuning(t)) || (rc = check_object_cov(b->scene, o))) return rc;
  if ((rc = batch_layout(b, t->tile_w))) return rc;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  if (!b->own_stream) HIP_TRY(hipStreamCreateWithFlags(&b->own_stream, hipStreamNonBlocking), "capture stream");
  HIP_TRY(hipStreamBeginCapture(b->own_stream, hipStreamCaptureModeThreadLocal), "begin capture");
  rc = batch_enqueue(b, o, t, b->own_stream);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->own_stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(e, "end capture");
  b->graph = g;
  HIP_TRY(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "graph instantiate");
  b->graph_tuning = *t;
  return CNQ_OK;
}

int cnq_batch_replay(cnq_batch* b, void* stream) {
  if (!b || !b->exec) return fail(CNQ_INVALID_ARG, "no captured graph (none yet, or the batch was laid out again)");
  HIP_TRY(hipGraphLaunch(b->exec, (hipStream_t)stream), "graph launch");
  return CNQ_OK;
}

int cnq_batch_autotune(cnq_batch* b, const cnq_options* o, void* stream, cnq_tuning* best, float* best_ms) {
  if (!b || !best) return fail(CNQ_INVALID_ARG, "null argument");
  int rc = check_options(o);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->scene->dev->ordinal);
  const uint32_t tiles[] = {16, 8, 32, 4, 64};
  const uint32_t per_cu[] = {0, 4, 8};
  const uint32_t waves[] = {2, 1};
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0), "event");
  HIP_TRY(hipEventCreate(&e1), "event");
  float best_t = 3.4e38f;
  cnq_tuning best_c{};
  // Two timed runs of a tuning after a warm one (which also lays the batch out).
  auto time_it = [&](const cnq_tuning& c, float& total) -> int {
    int r = cnq_batch_run(b, o, &c, st);
    total = 0.0f;
    for (int k = 0; k < 2 && !r; k++) {
      (void)hipEventRecord(e0, st);
      r = batch_enqueue(b, o, &c, st);
      (void)hipEventRecord(e1, st);
      (void)hipEventSynchronize(e1);
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      total += ms;
    }
    return r;
  };
  for (uint32_t tw : tiles) {
    for (uint32_t k : per_cu) {
      for (uint32_t wv : waves) {
        const cnq_tuning c{tw, k * (uint32_t)cus, wv, 64};
        float total = 0.0f;
        if ((rc = time_it(c, total))) break;
        if (total < best_t) {
          best_t = total;
          best_c = c;
        }
      }
      if (rc) break;
    }
    if (rc) break;
  }
  // The shallower ray stack, where it truncates no camera ray of the batch
  // (the scene's truncation counter does not move).
  if (!rc) {
    uint32_t before[4] = {0, 0, 0, 0}, after[4] = {0, 0, 0, 0};
    cnq_tuning c = best_c;
    c.ray_stack = 32;
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(before, b->scene->counters.p, sizeof(before), hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = hip_fail(e, "read counters");
    float total = 0.0f;
    if (!rc) rc = time_it(c, total);
    if (!rc && (e = hipMemcpy(after, b->scene->counters.p, sizeof(after), hipMemcpyDeviceToHost)) != hipSuccess)
      rc = hip_fail(e, "read counters");
    if (!rc && after[3] == before[3] && total < best_t) {
      best_t = total;
      best_c = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  *best = best_c;
  if (best_ms) *best_ms = best_t / 2.0f;
  return batch_layout(b, best_c.tile_w);
}

int cnq_batch_read(cnq_batch* b, uint32_t view, uint32_t* hits, float* visible, uint64_t* full, uint64_t* touched) {
  if (!b || view >= b->n_views) return fail(CNQ_INVALID_ARG, "no such view");
  HIP_TRY(hipSetDevice(b->scene->dev->ordinal), "hipSetDevice");
  HIP_TRY(hipDeviceSynchronize(), "sync");
  const cnq::ViewDesc& v = b->h_views[view];
  const uint64_t px = (uint64_t)v.w * v.h, q = px * v.n_lights;
  if (hits) HIP_TRY(hipMemcpy(hits, b->hits.p + 4 * v.recv_off, 16 * px, hipMemcpyDeviceToHost), "read hits");
  if (q && visible) HIP_TRY(hipMemcpy(visible, b->visible.p + v.out_off, 4 * q, hipMemcpyDeviceToHost), "read visible");
  if (q && full) HIP_TRY(hipMemcpy(full, b->full.p + v.out_off, 8 * q, hipMemcpyDeviceToHost), "read full");
  if (q && touched) HIP_TRY(hipMemcpy(touched, b->touched.p + v.out_off, 8 * q, hipMemcpyDeviceToHost), "read touched");
  return CNQ_OK;
}

int cnq_batch_set_lights(cnq_batch* b, const float* lights, uint32_t n_lights) {
  if (!b || (n_lights && !lights)) return fail(CNQ_INVALID_ARG, "null argument");
  if (n_lights != b->n_lights) return fail(CNQ_INVALID_ARG, "a batch keeps its light count");
  HIP_TRY(hipDeviceSynchronize(), "sync");
  HIP_TRY(hipMemcpy(b->lights.p, lights, 64 * (size_t)n_lights, hipMemcpyHostToDevice), "lights");
  return CNQ_OK;
}

}  // extern "C"

// ---- The lighting cache

struct cnq_cache {
  cnq_scene* scene;
  uint32_t n_blocks, n_slots;
  cnq_options opts;
  cnq::LightRect light;
  std::vector<uint8_t> dynamic;
  std::vector<cnq::host::Box> bounds;          // every mesh's current bounds (unpadded)
  std::vector<std::vector<float>> mesh_tris;   // dynamic meshes' current triangles (9 floats each)
  std::vector<uint32_t> node_span;             // per mesh: its triangle hierarchy's node count
  DevBuf<float> texels, planes, static_cov, irradiance;
  DevBuf<uint32_t> block_slot, slot_block, dirty, n_dirty, static_items;
  DevBuf<NodeC> static_nodes;
  DevBuf<uint64_t> bake_full, bake_touched;
  uint32_t rebake_grid = 0;

  void release() {
    texels.release(), planes.release(), static_cov.release(), irradiance.release();
    block_slot.release(), slot_block.release(), dirty.release(), n_dirty.release(), static_items.release();
    static_nodes.release(), bake_full.release(), bake_touched.release();
  }

  cnq::CacheArgs args() const {
    cnq::CacheArgs a;
    a.s = scene->args();
    a.static_nodes = static_nodes.p;
    a.static_items = static_items.p;
    a.texels = texels.p;
    a.planes = reinterpret_cast<const float4*>(planes.p);
    a.block_slot = block_slot.p;
    a.slot_block = slot_block.p;
    a.static_cov = static_cov.p;
    a.dirty = dirty.p;
    a.n_dirty = n_dirty.p;
    a.irradiance = irradiance.p;
    a.light = light;
    a.p = query_params(scene, &opts, 64);
    a.n_blocks = n_blocks;
    a.n_slots = n_slots;
    return a;
  }
};

namespace {

cnq::host::Box tri9_bounds(const float* t, size_t n) {
  cnq::host::Box b;
  for (int k = 0; k < 3; k++) {
    b.lo[k] = cnq::host::FMAX;
    b.hi[k] = -cnq::host::FMAX;
  }
  for (size_t v = 0; v < 3 * n; v++)
    for (int k = 0; k < 3; k++) {
      const float x = t[3 * v + k];
      b.lo[k] = x < b.lo[k] ? x : b.lo[k];
      b.hi[k] = x > b.hi[k] ? x : b.hi[k];
    }
  return b;
}

// The object hierarchy over the padded bounds of the meshes `keep` selects
// (`mesh_hierarchy_of`, hierarchy.rs:188-196), compacted.
void mesh_hierarchy_of(const std::vector<cnq::host::Box>& bounds, const uint8_t* keep, std::vector<NodeC>& out,
                       std::vector<uint32_t>& items) {
  std::vector<cnq::host::Box> padded(bounds.size());
  for (size_t m = 0; m < bounds.size(); m++) padded[m] = cnq::host::padded(bounds[m]);
  std::vector<cnq::host::NodeGpu> nodes;
  cnq::host::object_hierarchy(padded, nodes, items, keep);
  out.resize(nodes.size());
  for (size_t n = 0; n < nodes.size(); n++) out[n] = compact(reinterpret_cast<const uint32_t*>(&nodes[n]), true);
}

}  // namespace

extern "C" {

int cnq_block_planes(const float* texels, const float* blocks, uint32_t n_blocks, const float* light, float* planes) {
  if (n_blocks && (!texels || !blocks || !planes)) return fail(CNQ_INVALID_ARG, "null buffer");
  if (!light) return fail(CNQ_INVALID_ARG, "light is null");
  if (light[3] > 1.5f) return fail(CNQ_INVALID_ARG, "the lighting cache takes a rectangular light");
  cnq::host::block_planes(texels, blocks, n_blocks, light, planes);
  return CNQ_OK;
}

int cnq_cache_create(cnq_scene* s, const float* texels, const float* blocks, uint32_t n_blocks, const float* light,
                     const uint8_t* dynamic, const float* irradiance, const cnq_options* o, cnq_cache** out) {
  if (!s || !out || !light || !dynamic) return fail(CNQ_INVALID_ARG, "null argument");
  if (n_blocks && (!texels || !blocks)) return fail(CNQ_INVALID_ARG, "null buffer");
  int rc = check_options(o);
  if (rc) return rc;
  if (o->hierarchy != 2) return fail(CNQ_INVALID_ARG, "the lighting cache walks triangle hierarchies (hierarchy 2)");
  if (s->instanced) return fail(CNQ_INVALID_ARG, "the lighting cache takes a scene without transforms");
  if (light[3] > 1.5f) return fail(CNQ_INVALID_ARG, "the lighting cache takes a rectangular light");
  if (n_blocks > 0x3ffffffu) return fail(CNQ_CAPACITY, "too many lightmap blocks");
  HIP_TRY(hipSetDevice(s->dev->ordinal), "hipSetDevice");
  const uint32_t n_obj = s->n_objects;
  uint32_t n_dynamic = 0;
  for (uint32_t m = 0; m < n_obj; m++)
    if (dynamic[m] && s->h_ranges[4 * m + 1] > 0) n_dynamic++;
  if (n_dynamic > cnq::MAX_MOVING) return fail(CNQ_CAPACITY, "at most 16 dynamic meshes");

  // Meshes' triangles and hierarchy layout, from the device copy.
  std::vector<float> tris(12 * s->n_tris);
  std::vector<uint32_t> blas_items(s->n_blas_items);
  HIP_TRY(hipDeviceSynchronize(), "sync");
  if (s->n_tris) HIP_TRY(hipMemcpy(tris.data(), s->tris.p, 48 * s->n_tris, hipMemcpyDeviceToHost), "read tris");
  if (s->n_blas_items)
    HIP_TRY(hipMemcpy(blas_items.data(), s->blas_items.p, 4 * s->n_blas_items, hipMemcpyDeviceToHost), "read items");
  cnq_cache* c = new (std::nothrow) cnq_cache;
  if (!c) return fail(CNQ_OOM, "out of host memory");
  c->scene = s;
  c->n_blocks = n_blocks;
  c->opts = *o;
  std::memcpy(&c->light, light, 64);
  c->light.radiance = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
  c->dynamic.assign(dynamic, dynamic + n_obj);
  c->bounds.resize(n_obj);
  c->mesh_tris.resize(n_obj);
  c->node_span.assign(n_obj, 0);
  std::vector<uint32_t> roots;
  for (uint32_t m = 0; m < n_obj; m++)
    if (s->h_ranges[4 * m + 2] != NONE) roots.push_back(s->h_ranges[4 * m + 2]);
  std::sort(roots.begin(), roots.end());
  for (uint32_t m = 0; m < n_obj; m++) {
    const uint32_t first = s->h_ranges[4 * m], count = s->h_ranges[4 * m + 1], root = s->h_ranges[4 * m + 2];
    std::vector<float> t9(9 * (size_t)count);
    for (size_t v = 0; v < 3 * (size_t)count; v++)
      for (int k = 0; k < 3; k++) t9[3 * v + k] = tris[12 * (size_t)first + 4 * v + k];
    c->bounds[m] = tri9_bounds(t9.data(), count);
    if (!c->dynamic[m]) continue;
    if (root != NONE) {
      // A dynamic mesh's hierarchy is rebuilt in place: its items must be its
      // own triangles at its own offsets, as `flatten` lays them out.
      for (uint32_t k = first; k < first + count; k++)
        if (k >= blas_items.size() || blas_items[k] < first || blas_items[k] >= first + count) {
          delete c;
          return fail(CNQ_INVALID_ARG, "mesh " + std::to_string(m) + ": not laid out by flatten");
        }
      auto next = std::upper_bound(roots.begin(), roots.end(), root);
      c->node_span[m] = (next == roots.end() ? s->n_blas : *next) - root;
    }
    c->mesh_tris[m] = std::move(t9);
  }
  tris.clear();
  tris.shrink_to_fit();

  // Planes, slots, the static meshes' hierarchy.
  std::vector<float> planes(64 * (size_t)n_blocks), soa(64 * (size_t)n_blocks);
  cnq::host::block_planes(texels, blocks, n_blocks, light, planes.data());
  for (uint32_t b = 0; b < n_blocks; b++)
    for (uint32_t p = 0; p < cnq::BLOCK_PLANES; p++)
      std::memcpy(&soa[4 * ((size_t)p * n_blocks + b)], &planes[64 * (size_t)b + 4 * p], 16);
  std::vector<uint32_t> block_slot(n_blocks), slot_block;
  for (uint32_t b = 0; b < n_blocks; b++) {
    const float* p0 = &planes[64 * (size_t)b];
    const bool never = p0[0] == 0.0f && p0[1] == 0.0f && p0[2] == 0.0f && p0[3] == -1.0f;
    block_slot[b] = never ? NONE : (uint32_t)slot_block.size();
    if (!never) slot_block.push_back(b);
  }
  c->n_slots = (uint32_t)slot_block.size();
  std::vector<uint8_t> keep(n_obj);
  for (uint32_t m = 0; m < n_obj; m++) keep[m] = c->dynamic[m] ? 0 : 1;
  std::vector<NodeC> snodes;
  std::vector<uint32_t> sitems;
  mesh_hierarchy_of(c->bounds, keep.data(), snodes, sitems);

  hipError_t e = c->texels.upload(texels, 512 * (size_t)n_blocks);
  if (e == hipSuccess) e = c->planes.upload(soa.data(), soa.size());
  if (e == hipSuccess) e = c->block_slot.upload(block_slot.data(), block_slot.size());
  if (e == hipSuccess) e = c->slot_block.upload(slot_block.data(), slot_block.size());
  if (e == hipSuccess) e = c->static_cov.alloc((size_t)c->n_slots * 64 * 64);
  if (e == hipSuccess) e = c->irradiance.alloc(64 * (size_t)n_blocks);
  if (e == hipSuccess) e = c->dirty.alloc(n_blocks);
  if (e == hipSuccess) e = c->n_dirty.alloc(1);
  if (e == hipSuccess) e = c->static_nodes.upload(snodes.data(), snodes.size());
  if (e == hipSuccess) e = c->static_items.upload(sitems.data(), sitems.size());
  if (e == hipSuccess) e = hipMemset(c->n_dirty.p, 0, 4);
  if (e != hipSuccess) {
    c->release();
    delete c;
    return hip_fail(e, e == hipErrorOutOfMemory ? "cache buffers (out of device memory)" : "cache buffers");
  }
  int dev_cus = 0;
  (void)hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, s->dev->ordinal);
  c->rebake_grid = (uint32_t)std::max(1, dev_cus) * 8;
  e = cnq::launch_bake_static(c->args(), nullptr);
  if (e == hipSuccess && irradiance)
    e = hipMemcpy(c->irradiance.p, irradiance, 256 * (size_t)n_blocks, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    c->release();
    delete c;
    return hip_fail(e, "static coverage bake");
  }
  if (!irradiance && n_blocks) {
    // MaskConeKernel::bake: one query per texel through the whole scene.
    e = c->bake_full.alloc(64 * (size_t)n_blocks);
    if (e == hipSuccess) e = c->bake_touched.alloc(64 * (size_t)n_blocks);
    float l16[16];
    std::memcpy(l16, &c->light, 64);
    DevBuf<float> l;
    if (e == hipSuccess) e = l.upload(l16, 16);
    if (e != hipSuccess) {
      l.release();
      c->release();
      delete c;
      return hip_fail(e, "lightmap bake buffers");
    }
    rc = cnq_irradiance_device(s, c->texels.p, 64 * (uint64_t)n_blocks, l.p, 1, o, c->irradiance.p, c->bake_full.p,
                               c->bake_touched.p, nullptr, nullptr);
    e = hipDeviceSynchronize();
    l.release();
    c->bake_full.release();
    c->bake_touched.release();
    if (rc || e != hipSuccess) {
      c->release();
      delete c;
      return rc ? rc : hip_fail(e, "lightmap bake");
    }
  }
  HIP_TRY(hipDeviceSynchronize(), "cache set-up");
  *out = c;
  return CNQ_OK;
}

void cnq_cache_destroy(cnq_cache* c) {
  if (!c) return;
  (void)hipSetDevice(c->scene->dev->ordinal);
  (void)hipDeviceSynchronize();
  c->release();
  delete c;
}

int cnq_cache_update(cnq_cache* c, const uint32_t* meshes, uint32_t n, const float* tris, void* stream) {
  if (!c || (n && (!meshes || !tris))) return fail(CNQ_INVALID_ARG, "null argument");
  cnq_scene* s = c->scene;
  HIP_TRY(hipSetDevice(s->dev->ordinal), "hipSetDevice");
  hipStream_t st = (hipStream_t)stream;
  // Validate every change before touching anything.
  std::vector<size_t> offset(n);
  size_t off = 0;
  uint32_t changed = 0;
  for (uint32_t i = 0; i < n; i++) {
    const uint32_t m = meshes[i];
    if (m >= s->n_objects) return fail(CNQ_INVALID_ARG, "no mesh " + std::to_string(m));
    if (!c->dynamic[m]) return fail(CNQ_INVALID_ARG, "mesh " + std::to_string(m) + " has lightmap texels and cannot move");
    offset[i] = off;
    const size_t count = s->h_ranges[4 * m + 1];
    if (std::memcmp(tris + off, c->mesh_tris[m].data(), 36 * count) != 0) changed++;
    off += 9 * count;
  }
  if (changed > cnq::MAX_SWEPT) return fail(CNQ_CAPACITY, "at most 16 meshes per update");
  cnq::SweptBoxes sw{};
  sw.n = 0;
  std::vector<NodeC> cn;
  std::vector<float> leaf, t12;
  for (uint32_t i = 0; i < n; i++) {
    const uint32_t m = meshes[i];
    const uint32_t first = s->h_ranges[4 * m], count = s->h_ranges[4 * m + 1], root = s->h_ranges[4 * m + 2];
    const float* t9 = tris + offset[i];
    if (std::memcmp(t9, c->mesh_tris[m].data(), 36 * (size_t)count) == 0) continue;
    const cnq::host::Box old = c->bounds[m], now = tri9_bounds(t9, count);
    for (int k = 0; k < 3; k++) {
      (&sw.lo[sw.n].x)[k] = old.lo[k] < now.lo[k] ? old.lo[k] : now.lo[k];
      (&sw.hi[sw.n].x)[k] = old.hi[k] > now.hi[k] ? old.hi[k] : now.hi[k];
    }
    sw.lo[sw.n].w = sw.hi[sw.n].w = 0.0f;
    sw.n++;
    t12.assign(12 * (size_t)count, 0.0f);
    for (size_t v = 0; v < 3 * (size_t)count; v++)
      for (int k = 0; k < 3; k++) t12[4 * v + k] = t9[3 * v + k];
    HIP_TRY(hipMemcpyAsync(s->tris.p + 12 * (size_t)first, t12.data(), t12.size() * 4, hipMemcpyHostToDevice, st),
            "mesh triangles");
    if (root != NONE) {
      // Same count, so the same shape and offsets (live.rs:858-873).
      std::vector<cnq::host::NodeGpu> nodes;
      std::vector<uint32_t> items;
      std::vector<float> cov;
      cnq::host::triangle_hierarchy(t9, count, root, first, nodes, items, cov);
      if (nodes.size() != c->node_span[m]) return fail(CNQ_INVALID_ARG, "mesh hierarchy changed shape");
      cn.resize(nodes.size());
      for (size_t k = 0; k < nodes.size(); k++) cn[k] = compact(reinterpret_cast<const uint32_t*>(&nodes[k]), false);
      leaf.resize(9 * (size_t)count);
      for (size_t k = 0; k < count; k++) std::memcpy(&leaf[9 * k], t9 + 9 * (size_t)(items[k] - first), 36);
      HIP_TRY(hipMemcpyAsync(s->blas.p + root, cn.data(), cn.size() * sizeof(NodeC), hipMemcpyHostToDevice, st),
              "mesh nodes");
      HIP_TRY(hipMemcpyAsync(s->blas_cov.p + 27 * (size_t)root, cov.data(), cov.size() * 4, hipMemcpyHostToDevice, st),
              "mesh coverage");
      HIP_TRY(hipMemcpyAsync(s->blas_items.p + first, items.data(), items.size() * 4, hipMemcpyHostToDevice, st),
              "mesh items");
      HIP_TRY(hipMemcpyAsync(s->tri_leaf.p + 9 * (size_t)first, leaf.data(), leaf.size() * 4, hipMemcpyHostToDevice, st),
              "mesh leaves");
    }
    HIP_TRY(hipStreamSynchronize(st), "mesh upload");
    c->mesh_tris[m].assign(t9, t9 + 9 * (size_t)count);
    c->bounds[m] = now;
  }
  if (sw.n == 0) return CNQ_OK;
  // Median splits depend only on the mesh count: the layout is unchanged.
  std::vector<NodeC> nodes;
  std::vector<uint32_t> items;
  mesh_hierarchy_of(c->bounds, nullptr, nodes, items);
  if (nodes.size() != s->n_nodes || items.size() != s->n_items)
    return fail(CNQ_INVALID_ARG, "the scene's object hierarchy was not built by mesh_hierarchy");
  HIP_TRY(hipMemcpyAsync(s->nodes.p, nodes.data(), nodes.size() * sizeof(NodeC), hipMemcpyHostToDevice, st), "nodes");
  HIP_TRY(hipMemcpyAsync(s->items.p, items.data(), items.size() * 4, hipMemcpyHostToDevice, st), "items");
  HIP_TRY(hipStreamSynchronize(st), "object hierarchy upload");
  // Every dynamic mesh with triangles, by its padded bounds (live.rs:487-508).
  cnq::MovingMeshes mv{};
  mv.n = 0;
  for (uint32_t m = 0; m < s->n_objects; m++) {
    if (!c->dynamic[m] || s->h_ranges[4 * m + 1] == 0) continue;
    const cnq::host::Box p = cnq::host::padded(c->bounds[m]);
    mv.lo[mv.n] = make_float4(p.lo[0], p.lo[1], p.lo[2], 0.0f);
    mv.hi[mv.n] = make_float4(p.hi[0], p.hi[1], p.hi[2], 0.0f);
    mv.mesh[mv.n] = m;
    mv.n++;
  }
  const cnq::CacheArgs a = c->args();
  HIP_TRY(fill_u32(c->n_dirty.p, 0, 1, st), "clear marks");
  HIP_TRY(cnq::launch_mark_blocks(a, sw, st), "mark_blocks launch");
  HIP_TRY(cnq::launch_rebake(a, mv, std::min(c->rebake_grid, c->n_blocks), st), "rebake launch");
  return CNQ_OK;
}

int cnq_cache_read(cnq_cache* c, float* irradiance, uint32_t* blocks, uint32_t* n_blocks) {
  if (!c) return fail(CNQ_INVALID_ARG, "cache is null");
  HIP_TRY(hipSetDevice(c->scene->dev->ordinal), "hipSetDevice");
  HIP_TRY(hipDeviceSynchronize(), "sync");
  uint32_t nd = 0;
  HIP_TRY(hipMemcpy(&nd, c->n_dirty.p, 4, hipMemcpyDeviceToHost), "read marks");
  if (n_blocks) *n_blocks = nd;
  if (blocks && nd) HIP_TRY(hipMemcpy(blocks, c->dirty.p, 4 * (size_t)nd, hipMemcpyDeviceToHost), "read blocks");
  if (irradiance && c->n_blocks)
    HIP_TRY(hipMemcpy(irradiance, c->irradiance.p, 256 * (size_t)c->n_blocks, hipMemcpyDeviceToHost), "read texels");
  return CNQ_OK;
}

}  // extern "C"
