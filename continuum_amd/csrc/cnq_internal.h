// Internal to libcnq.so: the handles' definitions and the helpers the
// API's translation units share (status reporting, device buffers, launch
// set-up). Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/cnq.h"
#include "bvh_build.h"
#include "cnq_kernels.h"

namespace cnq {
namespace detail {

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

// Records `msg` as the calling thread's last error; returns `code`.
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define HIP_TRY(expr, what)              \
  do {                                   \
    hipError_t _e = (expr);              \
    if (_e != hipSuccess) return hip_fail(_e, what); \
  } while (0)

hipError_t fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s);
int check_options(const cnq_options* o);
uint64_t layout_view(ViewDesc& v, uint32_t tile_w, uint64_t item_off);

}  // namespace detail
}  // namespace cnq

struct cnq_device {
  int ordinal;
  uint32_t compute_units;  // hipDeviceProp_t::multiProcessorCount
  // This device's copy of the library's static face scratch (the address
  // hipGetSymbolAddress gives with the device current), for
  // cnq_sky_irradiance_device_noalloc.
  float* sky_scratch;
};

// A refit's schedule for some of the triangle hierarchies: their leaves,
// and their internal nodes by height (inner[level_off[h], level_off[h + 1])
// have height h + 1), built once on the host.
struct RefitPlan {
  cnq::detail::DevBuf<uint32_t> leaves, inner, level_off_d;
  std::vector<uint32_t> level_off;
  void release() {
    leaves.release(), inner.release(), level_off_d.release();
    level_off.clear();
  }
  uint64_t bytes() const { return leaves.bytes() + inner.bytes() + level_off_d.bytes(); }
};

// A launch's items in the order of their last times (longest first), for
// launching the same work again (a view's cone queries, a list's sky).
struct ItemOrder {
  bool bound = false, ready = false;
  void* stream = nullptr;  // the one stream the order is read and written on
  uint64_t key[4] = {0, 0, 0, 0};
  cnq::detail::DevBuf<uint32_t> cost, order;
  cnq::detail::DevBuf<uint8_t> last;  // each item's last time, as a power of two (0xff: none)
};

struct cnq_scene {
  cnq_device* dev;
  uint64_t n_tris;
  uint32_t n_objects, n_blas, n_nodes, n_items, n_xforms;
  uint32_t n_cov_rows = 0;  // rows of 27 coverage samples (>= n_blas)
  uint32_t bound_chunks = 1;  // parts of an object's vertices in the mesh-bounds pass
  uint64_t n_blas_items;
  uint32_t max_leaf = 0;  // the most triangles a triangle-hierarchy leaf held at upload
  bool count_work = false;
  uint8_t* query_flags = nullptr;  // cnq_scene_set_query_flags: caller-owned device bytes
  uint64_t n_query_flags = 0;
  cnq::detail::DevBuf<float> tris, tri_leaf, blas_cov, xforms, mesh_box;
  cnq::detail::DevBuf<float> normals;  // cnq_scene_set_normals: 9 per triangle (none: face normals)
  cnq::detail::DevBuf<uint32_t> blas_items, ranges, items, counters;
  RefitPlan refit;                  // every triangle hierarchy
  std::vector<uint32_t> h_child;    // per node: its first child, NONE for a leaf
  cnq::detail::DevBuf<uint64_t> work;
  cnq::detail::DevBuf<float> sky_faces;  // cnq_sky_irradiance_device's per-face values when the caller has none
  ItemOrder order;      // cnq_view_irradiance_device's
  ItemOrder sky_order;  // cnq_sky_irradiance_device's
  cnq::detail::DevBuf<cnq::NodeC> blas, nodes;
  std::vector<uint32_t> h_ranges;  // for host-side checks and object rebuilds
  // Moving objects (instanced scenes): the subtree's layout, each geometry's
  // object-space bounds, the movers' placements and which moved since the
  // last apply.
  bool has_movers = false;
  cnq::host::MoverTree movers;
  std::vector<cnq::host::Box> geometry_bounds;
  std::vector<cnq::host::Placement> mover_placements;
  std::vector<uint8_t> mover_dirty;
  std::vector<uint32_t> mover_of;  // per object: its index among the movers, or NONE
  std::vector<cnq::host::Box> mover_bounds;  // each mover's world box as last applied
  std::vector<cnq_moved> last_moves;         // the last apply_moves's moves
  std::map<uint32_t, RefitPlan> pose_plans;  // per posed geometry's root: its hierarchy's refit
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
    tris.release(), tri_leaf.release(), blas_cov.release(), xforms.release(), mesh_box.release(), normals.release();
    blas_items.release(), ranges.release(), items.release(), refit.release();
    counters.release(), work.release(), sky_faces.release(), order.cost.release(),
        order.order.release(), order.last.release(), sky_order.cost.release(), sky_order.order.release(),
        sky_order.last.release(), blas.release(), nodes.release();
    for (auto& kv : pose_plans) kv.second.release();
    pose_plans.clear();
  }
  uint64_t bytes() const {
    return tris.bytes() + tri_leaf.bytes() + blas_cov.bytes() + xforms.bytes() + mesh_box.bytes() +
           (normals.p ? normals.bytes() : 0) +
           blas_items.bytes() + ranges.bytes() + items.bytes() + refit.bytes() +
           counters.bytes() + work.bytes() + blas.bytes() + nodes.bytes();
  }
};

struct cnq_flat {
  cnq::host::Flat f;
  std::vector<float> xforms;      // instanced worlds: 16 per object
  uint32_t object_cov = 0xffffffffu;
  bool has_movers = false;
  cnq::host::MoverTree movers;
};

namespace cnq {
namespace detail {

QueryParams query_params(cnq_scene* s, const cnq_options* o, uint32_t tile_w);
ItemArgs item_args(cnq_scene* s, const float* receivers, const float* lights, float* vis, uint64_t* full,
                   uint64_t* touched, float* cov);
hipError_t launch_items_direct(ItemArgs a, uint64_t n_items, bool count_work, hipStream_t stream, uint32_t waves);
int check_object_cov(cnq_scene* s, const cnq_options* o);
// The refit schedule of the hierarchies under `roots` (child: per node its
// first child, NONE for a leaf), and one refit of the scene's triangle
// hierarchies by such a plan on `stream`.
int make_refit_plan(const std::vector<uint32_t>& child, const std::vector<uint32_t>& roots, RefitPlan& plan);
hipError_t run_refit(cnq_scene* s, const RefitPlan& plan, hipStream_t stream);

}  // namespace detail
}  // namespace cnq
