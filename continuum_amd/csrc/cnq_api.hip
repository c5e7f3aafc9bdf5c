// The C ABI (include/cnq.h): scene upload and validation, query launches,
// refit, diagnostics. Host code; the kernels live in cnq_cone.hip,
// cnq_primary.hip and cnq_refit.hip.
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/cnq.h"
#include "bvh_build.h"
#include "cnq_kernels.h"

using cnq::NodeC;
using cnq::NONE;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(CNQ_HIP_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr, what)              \
  do {                                   \
    hipError_t _e = (expr);              \
    if (_e != hipSuccess) return hip_fail(_e, what); \
  } while (0)

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  hipError_t alloc(size_t count) {
    n = count;
    return hipMalloc((void**)&p, (count ? count : 1) * sizeof(T));
  }
  hipError_t upload(const T* src, size_t count) {
    hipError_t e = alloc(count);
    if (e != hipSuccess) return e;
    if (count) e = hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice);
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  size_t bytes() const { return (n ? n : 1) * sizeof(T); }
};

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
// hierarchy.rs:751), one block per object.
__global__ __launch_bounds__(256) void mesh_bounds_kernel(const float* __restrict__ tris,
                                                          const uint32_t* __restrict__ ranges, float* out) {
  const uint32_t m = blockIdx.x;
  const uint32_t first = ranges[4 * m], count = ranges[4 * m + 1];
  float lo[3] = {3.40282347e38f, 3.40282347e38f, 3.40282347e38f};
  float hi[3] = {-3.40282347e38f, -3.40282347e38f, -3.40282347e38f};
  for (uint64_t v = threadIdx.x; v < 3ull * count; v += blockDim.x) {
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
    for (int k = 0; k < 6; k++) out[6 * m + k] = s[k][0];
}

}  // namespace

struct cnq_device {
  int ordinal;
};

struct cnq_scene {
  cnq_device* dev;
  uint64_t n_tris;
  uint32_t n_objects, n_blas, n_nodes, n_items, n_xforms;
  uint64_t n_blas_items;
  bool count_work = false;
  DevBuf<float> tris, tri_leaf, blas_cov, xforms, mesh_box;
  DevBuf<uint32_t> blas_items, ranges, items, parent, leaves, arrivals, counters;
  DevBuf<uint64_t> work;
  DevBuf<NodeC> blas, nodes;
  std::vector<uint32_t> h_ranges;  // for host-side checks and object rebuilds
  bool instanced = false;

  cnq::SceneArgs args() const {
    cnq::SceneArgs s;
    s.tris = tris.p;
    s.tri_leaf = tri_leaf.p;
    s.blas_items = blas_items.p;
    s.ranges = ranges.p;
    s.blas = blas.p;
    s.blas_cov = blas_cov.p;
    s.nodes = nodes.p;
    s.items = items.p;
    s.xforms = xforms.p;
    s.n_objects = n_objects;
    return s;
  }
  void release() {
    tris.release(), tri_leaf.release(), blas_cov.release(), xforms.release(), mesh_box.release();
    blas_items.release(), ranges.release(), items.release(), parent.release(), leaves.release();
    arrivals.release(), counters.release(), work.release(), blas.release(), nodes.release();
  }
  uint64_t bytes() const {
    return tris.bytes() + tri_leaf.bytes() + blas_cov.bytes() + xforms.bytes() + mesh_box.bytes() +
           blas_items.bytes() + ranges.bytes() + items.bytes() + parent.bytes() + leaves.bytes() + arrivals.bytes() +
           counters.bytes() + work.bytes() + blas.bytes() + nodes.bytes();
  }
};

struct cnq_flat {
  cnq::host::Flat f;
};

namespace {

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
  return CNQ_OK;
}

cnq::QueryArgs query_args(cnq_scene* s, const float* receivers, uint64_t n, const float* lights, uint32_t n_lights,
                          const cnq_options* o, float* vis, uint64_t* full, uint64_t* touched, float* cov) {
  cnq::QueryArgs a;
  a.s = s->args();
  a.receivers = receivers;
  a.lights = lights;
  a.n = n;
  a.q_begin = 0;
  a.q_end = n * (uint64_t)n_lights;
  a.hierarchy = o->hierarchy;
  a.light_cone_cap = o->light_cone_cap;
  a.wide_cone_cap = o->wide_cone_cap;
  a.profile_stage = o->profile_stage;
  a.object_cov = o->object_cov;
  a.prefilter = o->prefilter;
  a.wide_extent = o->wide_extent;
  a.out_visible = vis;
  a.out_full = full;
  a.out_touched = touched;
  a.out_cov = cov;
  a.counters = s->counters.p;
  a.work = s->count_work ? s->work.p : nullptr;
  return a;
}

int check_object_cov(cnq_scene* s, const cnq_options* o) {
  if (o->object_cov != NONE && (uint64_t)o->object_cov + s->n_nodes > s->n_blas)
    return fail(CNQ_INVALID_ARG, "object_cov + n_nodes exceeds the coverage samples");
  return CNQ_OK;
}

}  // namespace

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
  cnq_device* d = new (std::nothrow) cnq_device{ordinal};
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
  if (blas_cov) std::memcpy(blas_cov, f.blas_cov.data(), f.blas_cov.size() * 4);
  if (nodes) std::memcpy(nodes, f.nodes.data(), f.nodes.size() * 48);
  if (items) std::memcpy(items, f.items.data(), f.items.size() * 4);
  return CNQ_OK;
}

void cnq_flat_destroy(cnq_flat* flat) { delete flat; }

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
  s->n_nodes = d->n_nodes;
  s->n_items = d->n_items;
  s->n_xforms = d->n_xforms;
  s->h_ranges.assign(d->ranges, d->ranges + 4 * (size_t)d->n_objects);
  for (uint32_t m = 0; m < d->n_objects; m++)
    if (d->ranges[4 * m + 3] != NONE) s->instanced = true;
  std::vector<NodeC> blas(d->n_blas), nodes(d->n_nodes);
  std::vector<uint32_t> parent(d->n_blas, NONE), leaves;
  for (uint32_t n = 0; n < d->n_blas; n++) {
    const uint32_t* g = d->blas + 12 * (size_t)n;
    blas[n] = compact(g, false);
    if (g[9] > 0)
      leaves.push_back(n);
    else
      parent[g[8]] = parent[g[8] + 1] = n;
  }
  for (uint32_t n = 0; n < d->n_nodes; n++) nodes[n] = compact(d->nodes + 12 * (size_t)n, true);
  std::vector<float> tri_leaf(9 * d->n_blas_items);
  for (uint64_t k = 0; k < d->n_blas_items; k++) {
    const float* t = d->tris + 12 * (size_t)d->blas_items[k];
    float* o = tri_leaf.data() + 9 * k;
    o[0] = t[0], o[1] = t[1], o[2] = t[2], o[3] = t[4], o[4] = t[5], o[5] = t[6], o[6] = t[8], o[7] = t[9],
    o[8] = t[10];
  }
  std::vector<float> zero_cov;
  const float* cov = d->blas_cov;
  if (!cov) {
    zero_cov.assign(27 * (size_t)d->n_blas, 0.0f);
    cov = zero_cov.data();
  }
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = s->tris.upload(d->tris, 12 * d->n_tris);
  if (e == hipSuccess) e = s->tri_leaf.upload(tri_leaf.data(), tri_leaf.size());
  if (e == hipSuccess) e = s->blas_items.upload(d->blas_items, d->n_blas_items);
  if (e == hipSuccess) e = s->ranges.upload(d->ranges, 4 * (size_t)d->n_objects);
  if (e == hipSuccess) e = s->blas.upload(blas.data(), blas.size());
  if (e == hipSuccess) e = s->blas_cov.upload(cov, 27 * (size_t)d->n_blas);
  if (e == hipSuccess) e = s->nodes.upload(nodes.data(), nodes.size());
  if (e == hipSuccess) e = s->items.upload(d->items, d->n_items);
  if (e == hipSuccess) e = s->xforms.upload(d->xforms, 16 * (size_t)d->n_xforms);
  if (e == hipSuccess) e = s->parent.upload(parent.data(), parent.size());
  if (e == hipSuccess) e = s->leaves.upload(leaves.data(), leaves.size());
  if (e == hipSuccess) e = s->arrivals.alloc(d->n_blas);
  if (e == hipSuccess) e = s->mesh_box.alloc(6 * (size_t)d->n_objects);
  if (e == hipSuccess) e = s->counters.alloc(4);
  if (e == hipSuccess) e = s->work.alloc(6);
  if (e == hipSuccess) e = hipMemset(s->counters.p, 0, s->counters.bytes());
  if (e == hipSuccess) e = hipMemset(s->work.p, 0, s->work.bytes());
  if (e != hipSuccess) {
    s->release();
    delete s;
    return hip_fail(e, e == hipErrorOutOfMemory ? "scene upload (out of device memory)" : "scene upload");
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
  if (n > (1ull << 40) / n_lights) return fail(CNQ_CAPACITY, "too many queries");
  cnq::QueryArgs a = query_args(s, receivers, n, lights, n_lights, o, vis, full, touched, cov);
  // At most 2^31 - 1 waves per launch.
  const uint64_t per_launch = 64ull * 0x7fffffffull;
  for (uint64_t q0 = 0; q0 < a.q_end; q0 += per_launch) {
    cnq::QueryArgs b = a;
    b.q_begin = q0;
    b.q_end = q0 + per_launch < a.q_end ? q0 + per_launch : a.q_end;
    HIP_TRY(cnq::launch_cone_queries(b, s->count_work, (hipStream_t)stream), "cone query launch");
  }
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

int cnq_primary_device(cnq_scene* s, const float* camera, uint32_t width, uint32_t height, const float* lights,
                       uint32_t n_lights, const cnq_options* o, uint32_t* hits, float* receivers, void* stream) {
  if (!s) return fail(CNQ_INVALID_ARG, "scene is null");
  int rc = check_options(o);
  if (rc) return rc;
  if (width == 0 || height == 0) return CNQ_OK;
  if (!camera || !hits || !receivers || (n_lights && !lights)) return fail(CNQ_INVALID_ARG, "null buffer");
  cnq::PrimaryArgs a;
  a.s = s->args();
  a.camera = camera;
  a.lights = lights;
  a.n_lights = n_lights;
  a.width = width;
  a.height = height;
  a.hierarchy = o->hierarchy;
  a.nearest_first = o->nearest_first;
  a.hits = hits;
  a.receivers = receivers;
  a.counters = s->counters.p + 3;
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
  if (rc == CNQ_OK && n_lights) rc = cnq_irradiance_device(s, recv.p, px, l.p, n_lights, o, v.p, f.p, t.p, nullptr, nullptr);
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

int cnq_scene_refit_device(cnq_scene* s, const float* tris, void* stream) {
  if (!s || !tris) return fail(CNQ_INVALID_ARG, "null argument");
  if (s->instanced) return fail(CNQ_INVALID_ARG, "refit of instanced scenes goes through their geometries");
  hipStream_t st = (hipStream_t)stream;
  if (s->n_tris)
    HIP_TRY(hipMemcpyAsync(s->tris.p, tris, 48 * s->n_tris, hipMemcpyDeviceToDevice, st), "refit copy");
  HIP_TRY(fill_u32(s->arrivals.p, 0, s->n_blas, st), "refit clear");
  cnq::RefitArgs a;
  a.tris = s->tris.p;
  a.blas_items = s->blas_items.p;
  a.blas = s->blas.p;
  a.blas_cov = s->blas_cov.p;
  a.tri_leaf = s->tri_leaf.p;
  a.parent = s->parent.p;
  a.arrivals = s->arrivals.p;
  a.leaves = s->leaves.p;
  a.n_leaves = (uint32_t)s->leaves.n;
  HIP_TRY(cnq::launch_refit(a, st), "refit launch");
  // The object hierarchy over the new mesh bounds (live.rs:888-892 rebuilds
  // `mesh_hierarchy`): bounds on the device, the small tree on the host.
  hipLaunchKernelGGL(mesh_bounds_kernel, dim3(s->n_objects), dim3(256), 0, st, s->tris.p, s->ranges.p, s->mesh_box.p);
  HIP_TRY(hipGetLastError(), "mesh bounds launch");
  std::vector<float> mb(6 * (size_t)s->n_objects);
  HIP_TRY(hipMemcpyAsync(mb.data(), s->mesh_box.p, mb.size() * 4, hipMemcpyDeviceToHost, st), "mesh bounds");
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

}  // extern "C"
