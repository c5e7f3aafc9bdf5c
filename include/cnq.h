/* cnq.h: the C ABI of the MI355X cone-visibility library (libcnq.so).
 *
 * A drop-in for the GPU path behind the reference's `MaskConeKernel`
 * (crates/continuum-gpu/src/lib.rs:65) and the hierarchy it walks
 * (crates/continuum-gpu/src/hierarchy.rs). The reference has no FFI; these
 * are the entry points a Rust backend crate beside crates/continuum-gpu
 * would bind, with the reference's data layouts:
 *
 *   receiver  8 x f32   ReceiverGpu (lib.rs:56-59): point.xyz, _, normal.xyz, w
 *                       (w < 0: trace nothing; the reference packs w = 0)
 *   light    16 x f32   LightGpu (lib.rs:712-748): corner (w: 0 two-sided,
 *                       1 one-sided, 2 sphere), edge_u, edge_v, radiance
 *   node     12 x u32   NodeGpu (hierarchy.rs:70-76): lo vec4, hi vec4, info uvec4
 *   triangle 12 x f32   three vec4 vertices (hierarchy.rs:10-13)
 *   range     4 x u32   first triangle, count, hierarchy root, transform (hierarchy.rs:14-16)
 *   xform    16 x f32   Placement::rows (hierarchy.rs:274-286)
 *   camera   20 x f32   CameraUniform (lib.rs:3538-3553): eye, forward, right, up,
 *                       params (width, height, tan(fov/2), radiance)
 *
 * A query is one (receiver i, light l) pair at index l * n + i. Its results
 * are the irradiance per unit radiance (f32, as `hit_irradiance_rect`
 * returns it), the bins fully covered (`full`, bit b set where coverage[b]
 * >= 1) and the bins touched (`touched`, coverage[b] > 0), and optionally
 * the 64 coverage values.
 *
 * Conventions: every function returns a status (CNQ_OK = 0) and never
 * throws; cnq_last_error() describes the last failure on the calling
 * thread. Outputs are caller-allocated. A handle is used from one host
 * thread at a time. `*_device` functions take device pointers and a
 * hipStream_t (NULL: the null stream) and return once work is enqueued.
 */
#ifndef CNQ_H
#define CNQ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  CNQ_OK = 0,
  CNQ_NO_DEVICE = 1,
  CNQ_INVALID_ARG = 2,
  CNQ_CAPACITY = 3,
  CNQ_OOM = 4,
  CNQ_HIP_ERROR = 5
};

#define CNQ_NONE 0xffffffffu

typedef struct cnq_device cnq_device;
typedef struct cnq_scene cnq_scene;
typedef struct cnq_flat cnq_flat;

/* Kernel options: the reference's light-uniform fields and pipeline
 * overrides (lib.rs:763-801 `RectLight::uniform`; mask_cones.wgsl
 * WIDE_EXTENT :1047, LIGHT_CONE_CAP :1051, WIDE_CONE_CAP :1055). */
typedef struct cnq_options {
  uint32_t hierarchy;      /* Hierarchy (hierarchy.rs:727-741): 0 None, 1 Meshes, 2 Triangles */
  float prefilter;         /* set_prefilter (lib.rs:1507): bins, 0 = exact */
  float wide_extent;       /* set_wide_cones (lib.rs:1501): 8, or 1e30 when off */
  uint32_t light_cone_cap; /* 4096, or CNQ_NONE when off */
  uint32_t wide_cone_cap;  /* 512, or CNQ_NONE when off */
  uint32_t profile_stage;  /* set_profile_stage (lib.rs:1675): 0 off */
  uint32_t object_cov;     /* light.objects.x: first object node in blas_cov, CNQ_NONE: none */
  uint32_t nearest_first;  /* set_nearest_first (lib.rs:1640): camera rays' child order */
} cnq_options;

/* The reference's defaults (lib.rs:1215-1240): Triangles, exact, wide cones on. */
void cnq_options_default(cnq_options* opts);

const char* cnq_last_error(void);

/* MaskConeKernel::new (lib.rs:1155): CNQ_NO_DEVICE where the reference
 * returns None. */
int cnq_device_create(int ordinal, cnq_device** out);
void cnq_device_destroy(cnq_device* dev);

/* ---- Host hierarchy builder: `flatten` (hierarchy.rs:29-66) plus
 * `mesh_hierarchy` (hierarchy.rs:183-196) over meshes given as 9 floats per
 * triangle, counts[m] triangles each. */
int cnq_flat_build(const float* tris9, const uint32_t* counts, uint32_t n_meshes, cnq_flat** out);
/* sizes: n_tris, n_ranges, n_blas, n_blas_items, n_nodes, n_items */
int cnq_flat_sizes(const cnq_flat* flat, uint64_t sizes[6]);
int cnq_flat_copy(const cnq_flat* flat, float* tris, uint32_t* ranges, uint32_t* blas, uint32_t* blas_items,
                  float* blas_cov, uint32_t* nodes, uint32_t* items);
void cnq_flat_destroy(cnq_flat* flat);

/* An object placing one geometry (hierarchy.rs:262-301 `Placement`):
 * world = scale * rotation * p + translation. */
typedef struct cnq_placement {
  uint32_t geometry;
  float rotation[9];    /* row-major, orthonormal */
  float scale;
  float translation[3]; /* relative to the render origin */
  float clear;          /* one less the opacity of a blended object; 0: opaque */
} cnq_placement;

/* The instanced world: `flatten_instanced` (hierarchy.rs:464-522), the
 * object hierarchy over the objects' padded world bounds (lib.rs:1984)
 * and, with object_prefilter, the object hierarchy's coverage samples
 * (`object_coverage`, hierarchy.rs:661-700) after the triangle
 * hierarchies' in blas_cov (lib.rs:2271-2278). cnq_flat_sizes/copy give the
 * arrays as for flatten (blas_cov: the triangle hierarchies' rows); then: */
int cnq_flat_build_instanced(const float* tris9, const uint32_t* counts, uint32_t n_geometries,
                             const cnq_placement* objects, uint64_t n_objects, int object_prefilter, cnq_flat** out);
/* The same with `movable` objects (lib.rs:1986-1993): the object hierarchy
 * is `object_hierarchy_with_movers` (hierarchy.rs:331-372): node 1 roots
 * the static objects, node 2 the movers' subtree (rebuilt as they move). */
int cnq_flat_build_instanced_movers(const float* tris9, const uint32_t* counts, uint32_t n_geometries,
                                    const cnq_placement* objects, uint64_t n_objects, const uint32_t* movable,
                                    uint32_t n_movable, int object_prefilter, cnq_flat** out);

/* hierarchy.rs:308-325 `MoverTree`. */
typedef struct cnq_mover_tree {
  uint32_t fixed_root[12];  /* the static objects' root (node 1), NodeGpu */
  uint32_t first_node;      /* the movers' nodes after node 2 */
  uint32_t nodes;
  uint32_t first_item;      /* the movers' items */
  uint32_t n_objects;
  const uint32_t* objects;  /* the movable objects, in subtree order */
} cnq_mover_tree;

/* The flat's mover tree (`objects` points into the flat). */
int cnq_flat_mover_tree(const cnq_flat* flat, cnq_mover_tree* out);
/* sizes: n_xforms, coverage rows in blas_cov (with the object hierarchy's),
 * object_cov (the object hierarchy's first row; CNQ_NONE: none). */
int cnq_flat_instanced_sizes(const cnq_flat* flat, uint64_t sizes[3]);
int cnq_flat_copy_instanced(const cnq_flat* flat, float* xforms, float* blas_cov);

/* ---- A scene resident in HBM, from the reference's flattened arrays
 * (lib.rs:3361-3440 `run`'s buffers; lib.rs:3002 `render_buffers_with`).
 * Every index is checked on the host: children follow their parents,
 * leaves and ranges stay inside their arrays. */
typedef struct cnq_scene_desc {
  const float* tris;          uint64_t n_tris;       /* 12 f32 per triangle */
  const uint32_t* ranges;     uint32_t n_objects;    /* 4 u32 per object */
  const uint32_t* blas;       uint32_t n_blas;       /* NodeGpu */
  const uint32_t* blas_items; uint64_t n_blas_items;
  const float* blas_cov;                             /* 27 per blas node (NULL: zeros) */
  const uint32_t* nodes;      uint32_t n_nodes;      /* object hierarchy, NodeGpu */
  const uint32_t* items;      uint32_t n_items;
  const float* xforms;        uint32_t n_xforms;     /* 16 f32 per transform */
  uint32_t n_blas_cov;                               /* rows in blas_cov (0: n_blas) */
} cnq_scene_desc;

int cnq_scene_create(cnq_device* dev, const cnq_scene_desc* desc, cnq_scene** out);
void cnq_scene_destroy(cnq_scene* scene);
/* Bytes of HBM the scene holds. */
uint64_t cnq_scene_bytes(const cnq_scene* scene);

/* ---- Moving objects of an instanced scene (the reference's
 * `Renderer::set_transform`, renderer.rs:442, applied by `apply_moves`,
 * instanced.rs:1226-1410). set_movers gives the scene its mover tree, each
 * geometry's object-space bounds (6 floats: lo, hi) and the movers'
 * current placements (tree order). set_transform gives a mover a new
 * rotation (row-major) and translation; apply_moves writes the moved
 * objects' transforms and rebuilds the movers' subtree (`mover_subtree`,
 * `union_root`, hierarchy.rs:374-411) on `stream`. */
int cnq_scene_set_movers(cnq_scene* scene, const cnq_mover_tree* tree, const float* geometry_bounds,
                         uint32_t n_geometries, const cnq_placement* placements);
int cnq_scene_set_transform(cnq_scene* scene, uint32_t object, const float* rotation, const float* translation);
int cnq_scene_apply_moves(cnq_scene* scene, void* stream);
/* A new pose of a deforming geometry (instanced.rs:1154-1223 `apply_poses`):
 * its `n_tris` triangles (9 floats each, the geometry's count and order)
 * replace its vertices, its triangle hierarchy is refit on the GPU (same
 * topology and leaf order, `refit_triangle_hierarchy`, hierarchy.rs:579),
 * its bounds are recomputed (`mesh_bounds`), and every object placing it
 * (all must be movable) is marked moved, so the next cnq_scene_apply_moves
 * rebuilds the movers' subtree over their new boxes. */
int cnq_scene_set_pose(cnq_scene* scene, uint32_t geometry, const float* tris9, uint32_t n_tris, void* stream);

/* ---- Queries. MaskConeKernel::irradiance_with (lib.rs:1718) and
 * `coverage` (lib.rs:3289): host buffers in and out, synchronous. out_cov
 * (64 per query) may be NULL. */
int cnq_irradiance(cnq_scene* scene, const float* receivers, uint64_t n, const float* lights, uint32_t n_lights,
                   const cnq_options* opts, float* out_visible, uint64_t* out_full, uint64_t* out_touched,
                   float* out_cov);

/* The same on device buffers, enqueued on `stream` (the WGSL `main` entry,
 * mask_cones.wgsl:1513, dispatched as lib.rs:3497-3512 does). */
int cnq_irradiance_device(cnq_scene* scene, const float* receivers, uint64_t n, const float* lights,
                          uint32_t n_lights, const cnq_options* opts, float* out_visible, uint64_t* out_full,
                          uint64_t* out_touched, float* out_cov, void* stream);

/* ---- The sky light: a uniform sky through six cube-face mask cones
 * (mask_cones.wgsl :1323-1510, `MaskConeKernel::set_sky` lib.rs:1520). */
typedef struct cnq_sky_options {
  uint32_t max_combine;  /* params.w bit 1 (`sky_max`, lib.rs:103): meshes combine by the maximum (a
                            diagnostic); 0: as independent occluders, 1 - (1 - a)(1 - b) (WGSL :820) */
  uint32_t horizon_cull; /* `sky_horizon_cull` (lib.rs:105): walks skip geometry wholly below the receiver's
                            horizon (default 1) */
  float prefilter_scale; /* `sky_prefilter` (lib.rs:104): the sky cones' prefilter is cnq_options.prefilter
                            times this (default 1) */
} cnq_sky_options;

void cnq_sky_options_default(cnq_sky_options* sky);

/* `sky_irradiance` (WGSL :1411-1426) at each receiver (ReceiverGpu, 32 B; a
 * negative normal.w traces nothing): out[i] the irradiance per unit sky
 * radiance, the six faces' light summed in order (+x, -x, +y, -y, +z, -z;
 * each face's `right` axis the next world axis). out_faces (nullable, 6 x n:
 * face f of receiver i at f * n + i) and out_full (nullable, 6 x n: the
 * face's bins with coverage 1) give the faces apart. Device buffers,
 * enqueued on `stream`; without out_faces the scene keeps one 6 x n buffer
 * for the faces, shared by every such call: calls without out_faces on
 * different streams must not overlap, and a launch captured into a graph
 * must pass out_faces (CNQ_INVALID_ARG otherwise). */
int cnq_sky_irradiance_device(cnq_scene* scene, const float* receivers, uint64_t n, const cnq_options* opts,
                              const cnq_sky_options* sky, float* out, float* out_faces, uint64_t* out_full,
                              void* stream);
/* The same over a width x height view's receivers (pixel y * width + x;
 * n = width * height in the layouts above), each wave taking a tile_w x
 * 64 / tile_w patch of pixels (tile_w a power of two up to 64) instead of
 * 64 consecutive receivers: the cones of a patch start close together, so
 * their walks share more nodes. Results equal the list launch's bit for bit
 * (sky walks have no node cap and keep a fixed child order). */
int cnq_sky_view_irradiance_device(cnq_scene* scene, const float* receivers, uint32_t width, uint32_t height,
                                   uint32_t tile_w, const cnq_options* opts, const cnq_sky_options* sky, float* out,
                                   float* out_faces, uint64_t* out_full, void* stream);
/* The same on host buffers (copied in and out; returns when done). */
int cnq_sky_irradiance(cnq_scene* scene, const float* receivers, uint64_t n, const cnq_options* opts,
                       const cnq_sky_options* sky, float* out, float* out_faces, uint64_t* out_full);

/* No counterpart in the reference: for callers that cannot allocate (code
 * under graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * cnq_sky_irradiance_device without out_faces and without allocating: the
 * per-face values go to a static array of the library's (6 x
 * CNQ_SKY_NOALLOC_MAX floats, 24 MiB, one copy per GPU, held until the
 * process ends), items run in natural order (no order buffers), and nothing
 * is allocated or freed. At most CNQ_SKY_NOALLOC_MAX receivers a call
 * (CNQ_CAPACITY beyond). Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (other streams, other scenes, other threads)
 * race. Results equal cnq_sky_irradiance_device's bit for bit. */
#define CNQ_SKY_NOALLOC_MAX (1u << 20)
int cnq_sky_irradiance_device_noalloc(cnq_scene* scene, const float* receivers, uint64_t n, const cnq_options* opts,
                                      const cnq_sky_options* sky, float* out, uint64_t* out_full, void* stream);

/* The same over a width x height view's receivers (pixel y * width + x),
 * each wave taking a tile_w x (64 / tile_w) patch of pixels (tile_w a power
 * of two up to 64): the cone part of the WGSL `render` entry
 * (mask_cones.wgsl:2294, 8 x 8 workgroups, lib.rs:3126-3134). Query
 * (light l, pixel p) is stored at l * width * height + p. */
int cnq_view_irradiance_device(cnq_scene* scene, const float* receivers, uint32_t width, uint32_t height,
                               const float* lights, uint32_t n_lights, const cnq_options* opts, uint32_t tile_w,
                               float* out_visible, uint64_t* out_full, uint64_t* out_touched, void* stream);
/* (Both this and cnq_sky_irradiance_device record each work item's time
 * and launch the same view again, on the same stream, heaviest items first;
 * the scene keeps that order for one view (receivers buffer and sizes) and
 * binds it to the first stream it was used on, where launches see it in
 * stream order. Launches on other streams, and launches captured into a
 * graph, run in natural order and record no times. The results do not
 * depend on the order.) */

/* Camera rays for a width x height view (the WGSL `visibility` entry,
 * mask_cones.wgsl:2227, with `camera_hit` :2222): per pixel, hits (mesh,
 * triangle, distance bits, seen light + 1) and receivers (the shading point
 * offset by `surface_offset`, the face normal toward the camera or, with
 * cnq_scene_set_normals, the interpolated shading normal; w = -1 where the
 * pixel sees the sky or a light). Device buffers. */
int cnq_primary_device(cnq_scene* scene, const float* camera, uint32_t width, uint32_t height, const float* lights,
                       uint32_t n_lights, const cnq_options* opts, uint32_t* hits, float* receivers, void* stream);

/* Corner shading normals (the reference's per-triangle `normals`, vec4 x 3
 * per triangle in the scene's triangle order, lib.rs:3002
 * `render_buffers_with`; `Renderer::set_vertices(g, tris, normals)`,
 * renderer.rs:430): 9 floats per triangle, n_tris the scene's count; NULL
 * drops them. With normals, every camera pass (cnq_primary_device,
 * cnq_render_view, batches, served views) gives its receivers the
 * interpolated shading normal (`shading_normal`, mask_cones.wgsl
 * :2047-2058: the corners' normals at the hit's barycentrics, the face's
 * where they sum to nothing, turned to the world, on the camera's side);
 * the shading point stays offset along the face normal. Synchronous. */
int cnq_scene_set_normals(cnq_scene* scene, const float* normals, uint64_t n_tris);

/* The plain per-view launch (MaskConeKernel::render_instanced_lights,
 * lib.rs:1857, without shading): camera rays, then one query per pixel and
 * light. Host buffers: hits 4 per pixel, the rest one per query. */
int cnq_render_view(cnq_scene* scene, const float* camera, uint32_t width, uint32_t height, const float* lights,
                    uint32_t n_lights, const cnq_options* opts, uint32_t* out_hits, float* out_visible,
                    uint64_t* out_full, uint64_t* out_touched);

/* ---- Refit (hierarchy.rs:579 `refit_triangle_hierarchy`, applied to
 * every mesh with a hierarchy; instanced.rs:1154 `apply_poses`): new
 * triangle positions (12 f32 per triangle, same count and order), on the
 * device in one bottom-up pass. */
int cnq_scene_refit_device(cnq_scene* scene, const float* tris, void* stream);
int cnq_scene_refit(cnq_scene* scene, const float* tris);
/* The scene's triangle hierarchy as it is now (NodeGpu layout) and its
 * coverage samples (27 per node), for reading back a refit. */
int cnq_scene_read_blas(cnq_scene* scene, uint32_t* blas, float* blas_cov);

/* ---- Diagnostics: counters accumulated by queries since the last reset:
 * [0] capped walks, [1] walks deeper than the reference's 48-entry stack,
 * [2] walks this build truncated (cones, or camera rays deeper than 48);
 * work, summed over queries when count_work is on (KernelOptions::
 * count_work, lib.rs:523-537; WGSL :655-665): object nodes tested,
 * triangle-hierarchy nodes tested, triangles tested, triangles that
 * reached the rasterizer, meshes entered, triangle-hierarchy nodes
 * splatted. */
int cnq_scene_counters(cnq_scene* scene, int reset, uint32_t counters[3], uint64_t work[6]);
int cnq_scene_set_count_work(cnq_scene* scene, int on);
/* Per-query walk flags for the queries launched after this call (NULL or
 * n = 0: none): a device buffer of n bytes; every cone-query launch on the
 * scene (cnq_irradiance[_device], cnq_view_irradiance_device,
 * cnq_render_view, batches: the batch's concatenated output, view after
 * view) writes, for each query whose output index is below n, bit 0 the walk
 * stopped at its cone cap (light_cone_cap / wide_cone_cap; the walk's
 * stop in `scene_coverage_if`, WGSL :1160-1227), bit 1 it went deeper than the reference's 48-entry stack, bit 2
 * this build truncated it. A capped walk's coverage depends on the order
 * nodes were visited in; the flags name exactly those queries. The buffer
 * must outlive the launches that write it. Served views write none. */
int cnq_scene_set_query_flags(cnq_scene* scene, uint8_t* flags, uint64_t n);

/* ---- Batches: many views of uneven size and light count in one camera
 * pass and one cone launch (no counterpart in the reference, which renders
 * one view per `render_instanced_lights` call, lib.rs:1857). Each view has
 * its camera, size and a range of the batch's lights; results are read per
 * view in the layout of cnq_render_view. */
typedef struct cnq_view {
  float camera[20];
  uint32_t width, height;
  uint32_t light_first, light_count;
} cnq_view;

/* How the passes are shaped: the pixel tile a cone wave takes (tile_w x
 * 64 / tile_w); if non-zero, a persistent grid of that many one-wave blocks
 * that take work items from a counter (else one block per item); the cone
 * walk's waves in flight per SIMD (2, the most its registers and LDS allow,
 * or 1: half the waves resident, half the lanes' private memory in the
 * caches; 0 means 2); and the camera rays' stack depth (64 entries, or 32:
 * half the LDS, more waves in flight; a ray deeper than the depth stops and
 * is counted as truncated; 0 means 64). Results do not depend on the
 * tuning, except that a ray stack too shallow for the scene truncates rays
 * (the autotuner takes 32 only where it truncates none). */
typedef struct cnq_tuning {
  uint32_t tile_w;
  uint32_t persistent_blocks;
  uint32_t waves_per_simd;
  uint32_t ray_stack;
} cnq_tuning;

typedef struct cnq_batch cnq_batch;

void cnq_tuning_default(cnq_tuning* t);
int cnq_batch_create(cnq_scene* scene, const cnq_view* views, uint32_t n_views, const float* lights,
                     uint32_t n_lights, cnq_batch** out);
void cnq_batch_destroy(cnq_batch* batch);
uint64_t cnq_batch_queries(const cnq_batch* batch);
/* Enqueue the batch's camera and cone passes on `stream`. */
int cnq_batch_run(cnq_batch* batch, const cnq_options* opts, const cnq_tuning* tuning, void* stream);
/* Record both passes as a hipGraph (the counter reset is a kernel, not a
 * memset), then replay it on any stream. */
int cnq_batch_capture(cnq_batch* batch, const cnq_options* opts, const cnq_tuning* tuning);
int cnq_batch_replay(cnq_batch* batch, void* stream);
/* Time every candidate tuning on this batch (tile shapes x persistent grid
 * sizes x waves per SIMD, then the shallower ray stack if it truncates no
 * ray) and return the fastest; results do not depend on the tuning. */
int cnq_batch_autotune(cnq_batch* batch, const cnq_options* opts, void* stream, cnq_tuning* best, float* best_ms);
/* Synchronize and copy one view's hits and query results. */
int cnq_batch_read(cnq_batch* batch, uint32_t view, uint32_t* hits, float* out_visible, uint64_t* out_full,
                   uint64_t* out_touched);
/* Replace the batch's lights (same count): a light edit. */
int cnq_batch_set_lights(cnq_batch* batch, const float* lights, uint32_t n_lights);

/* ---- Serving: a request loop over one resident scene (no counterpart in
 * the reference, which renders on the caller's thread; this is the
 * continuous-batching server its Renderer would sit behind). View requests
 * (a camera, a size, and 1..max_view_lights indices into the server's light
 * table) are admitted in arrival order (the express slot, when configured,
 * takes small views first): a view is admitted when its results fit in the
 * free HBM (48 B per pixel for shading points and hits, 20 B per query) and a
 * launch slot is free; admitted views go out together as one camera pass and
 * cone launches on the slot's stream, and launches on different slots
 * overlap on the device. A light edit applies to the table at once and so to
 * the views submitted after it. A view's memory stays held until
 * cnq_server_release. Submit, wait, read and release may be called from any
 * host thread. */
typedef struct cnq_server cnq_server;

typedef struct cnq_server_config {
  uint64_t pixel_capacity;       /* shading points of admitted views held at once */
  uint64_t query_capacity;       /* queries of admitted views held at once */
  uint32_t max_lights;           /* size of the light table */
  uint32_t max_view_lights;      /* lights per view */
  uint32_t max_views_per_launch; /* views batched into one launch */
  uint32_t slots;                /* launches in flight (one stream each) */
  cnq_tuning tuning;             /* cone launch shape */
  uint64_t launch_queries;       /* cone launches carry at most this many queries (whole
                                    64-query tiles); the rest of a launch's views follow as
                                    further launches on its stream, so other slots' launches
                                    reach the device in between (0: one launch) */
  uint64_t express_queries;      /* the last slot takes only views of at most this many
                                    queries, ahead of larger views that arrived earlier; the
                                    other slots admit in arrival order (0: no express slot;
                                    needs slots >= 2) */
} cnq_server_config;

enum { CNQ_REQ_QUEUED = 0, CNQ_REQ_ADMITTED = 1, CNQ_REQ_DONE = 2, CNQ_REQ_RELEASED = 3 };

typedef struct cnq_request_info {
  int32_t state;
  uint32_t width, height, n_lights;
  uint64_t launch;       /* the launch that carried it */
  double queued_ms;      /* submit to admission */
  double latency_ms;     /* submit to results readable */
  double admitted_at_ms; /* admission and completion, from the server's creation */
  double done_at_ms;
} cnq_request_info;

void cnq_server_config_default(cnq_server_config* cfg);
int cnq_server_create(cnq_scene* scene, const cnq_server_config* cfg, const cnq_options* opts, const float* lights,
                      uint32_t n_lights, cnq_server** out);
/* Waits for in-flight launches, drops queued requests, frees everything. */
void cnq_server_destroy(cnq_server* server);
int cnq_server_submit_view(cnq_server* server, const float* camera, uint32_t width, uint32_t height,
                           const uint32_t* lights, uint32_t n_lights, uint64_t* id);
int cnq_server_submit_light(cnq_server* server, uint32_t index, const float* light, uint64_t* id);
/* Block until request `id` is done or `timeout_ms` passes (*done = 0). */
int cnq_server_wait(cnq_server* server, uint64_t id, double timeout_ms, int* done);
int cnq_server_info(cnq_server* server, uint64_t id, cnq_request_info* info);
/* A done view's hits (4 u32 per pixel) and queries (light-major, as
 * cnq_render_view); any pointer may be NULL. */
int cnq_server_read(cnq_server* server, uint64_t id, uint32_t* hits, float* out_visible, uint64_t* out_full,
                    uint64_t* out_touched);
/* Free a done request's memory (admits more); its results are gone. */
int cnq_server_release(cnq_server* server, uint64_t id);
/* [views done, edits done, launches, queries done, pixels held, queries held] */
int cnq_server_counts(cnq_server* server, uint64_t counts[6]);

/* ---- The lighting cache: a lightmap's direct irradiance from one rect
 * light, kept current as meshes move (the reference's `LiveRenderer`,
 * crates/continuum-gpu/src/live.rs, over a `Lightmap`, lightmap.rs).
 * Texels come in blocks of 8 x 8: 64 consecutive receivers per block
 * (Lightmap::receivers; texels on no triangle have a zero normal), with
 * each block's bounds (Lightmap::blocks, 6 floats: lo, hi). Meshes flagged
 * dynamic may move; the others hold the texels. */
#define CNQ_BLOCK_PLANES 16 /* lightmap.rs:83 BLOCK_PLANES */
#define CNQ_MAX_SWEPT 16    /* live.rs:220 MAX_SWEPT: changed meshes per update */
#define CNQ_MAX_MOVING 16   /* live.rs:175 MAX_MOVING: dynamic meshes */

typedef struct cnq_cache cnq_cache;

/* Lightmap::block_planes (lightmap.rs:256-333): 16 planes (n.xyz, w;
 * inside n.p <= w) per block enclosing its texels' cones toward the rect
 * `light`. Host only. */
int cnq_block_planes(const float* texels, const float* blocks, uint32_t n_blocks, const float* light,
                     float* planes);

/* LiveRenderer::new's lightmap set-up (live.rs:515-720): block planes,
 * static-coverage slots, the static meshes' hierarchy (`mesh_hierarchy_of`,
 * hierarchy.rs:188) and the `bake_static` pass (WGSL :2468). `irradiance`
 * (64 per block) is the baked lightmap (Lightmap::irradiance); NULL bakes
 * it here as MaskConeKernel::bake does (lib.rs:2961). The scene must be
 * flattened by the reference's `flatten` (no transforms), `opts.hierarchy`
 * Triangles; the cache keeps `scene`, which must outlive it. */
int cnq_cache_create(cnq_scene* scene, const float* texels, const float* blocks, uint32_t n_blocks,
                     const float* light, const uint8_t* dynamic, const float* irradiance, const cnq_options* opts,
                     cnq_cache** out);
void cnq_cache_destroy(cnq_cache* cache);
/* LiveRenderer::update_meshes (live.rs:795-900): `n` dynamic meshes take
 * new triangles (9 floats each, the mesh's own count, concatenated in the
 * order of `meshes`). Each changed mesh's hierarchy is rebuilt in place and
 * the object hierarchy over all meshes rebuilt; then `mark_blocks` (WGSL
 * :2430) marks the blocks its swept bounds can reach and `rebake` (WGSL
 * :2514) recomputes their texels, enqueued on `stream`. */
int cnq_cache_update(cnq_cache* cache, const uint32_t* meshes, uint32_t n, const float* tris, void* stream);
/* Synchronize and copy the texels' irradiance (64 per block, nullable), the
 * blocks the last update re-baked (nullable; any order) and their count. */
int cnq_cache_read(cnq_cache* cache, float* irradiance, uint32_t* blocks, uint32_t* n_blocks);

/* ---- The view cache's pages (the reference's `VirtualCache`,
 * crates/continuum-gpu/src/vcache.rs, and its passes in mask_cones.wgsl).
 * A page is 9 x 9 texels (8 and a border) of one chart of one object at one
 * mip, keyed (object, chart | mip << 24, page x | page y << 16); texel
 * (x, y) of a page holds, per light, the irradiance per unit radiance (or,
 * with `light_ratio`, the light's shadowed share of its unshadowed light)
 * through one mask cone from the texel's place on the surface, then one
 * sky channel (zero here: no sky). Resident pages are found through an
 * open-addressing page table of (key, slot) entries. When objects move,
 * the stale passes find the texels whose cones the moved boxes can reach
 * and re-bake them (pages in view) or drop their pages (out of view). */
#define CNQ_PAGE 8          /* vcache.rs:35 PAGE */
#define CNQ_PAGE_TEXELS 81  /* vcache.rs:37 PAGE_TEXELS */
#define CNQ_VC_MAX_LIGHTS 16 /* lib.rs:751 MAX_LIGHTS */

typedef struct cnq_vcache cnq_vcache;

/* The scene's surfaces (continuum::mesh::Surface, crates/continuum/src/
 * mesh.rs:20-39), per triangle in the scene's order (geometries, then their
 * triangles): its global chart, its corners' chart coordinates (6 floats,
 * in [0, 1]^2) and shading normals (9 floats); per global chart: its size
 * along u and v (object units) and its cell grid. */
typedef struct {
  const uint32_t* tri_chart;
  const float* tri_uv;
  const float* normals;
  uint32_t n_charts;
  const float* chart_size;
  const uint32_t* chart_grid;
} cnq_vcache_surfaces;

typedef struct {
  uint32_t pages;          /* resident pages (CacheOptions::pages) */
  uint32_t append_pages;   /* pages the stale pass may add to a frame's bake (vcache.rs append_pages) */
  uint32_t candidates;     /* candidate pages a frame's stale pass may list */
  float base_texel;        /* a mip-0 texel's size in object units (CacheOptions::base_texel) */
  uint32_t light_ratio;    /* 1: texels keep the shadowed share (CacheOptions::light_ratio) */
  uint32_t visible_frames; /* a page used within this many frames is in view (re-baked when stale) */
} cnq_vcache_config;

/* An object moved since the last frame (vcache.rs:62-76 `Moved`, rigid):
 * its world box before and after. */
typedef struct {
  uint32_t object;
  float old_lo[3], old_hi[3], new_lo[3], new_hi[3];
} cnq_moved;

/* The moves the scene's last cnq_scene_apply_moves made (instanced.rs:
 * 1284-1306: each moved object's world box before and after, `world_bounds`
 * of its placement), for cnq_vcache_set_moved. */
int cnq_scene_last_moves(cnq_scene* scene, cnq_moved* out, uint32_t cap, uint32_t* n);

void cnq_vcache_config_default(cnq_vcache_config* cfg);
/* The view cache over `scene` (which must outlive it) with `n_lights`
 * static lights (LightGpu rows): chart cell tables (vcache.rs:704
 * `chart_tables`), an empty page table (vcache.rs:1497, 4 entries per page
 * rounded up to a power of two, at least 1,024), the texel pool. */
int cnq_vcache_create(cnq_scene* scene, const cnq_vcache_surfaces* surfaces, const float* lights, uint32_t n_lights,
                      const cnq_vcache_config* cfg, const cnq_options* opts, cnq_vcache** out);
void cnq_vcache_destroy(cnq_vcache* cache);
/* Texels along u and v of `chart` on `object` at `mip` (vcache.rs:2104
 * `grid`, WGSL :3181 `vc_grid`). */
int cnq_vcache_grid(cnq_vcache* cache, uint32_t object, uint32_t chart, uint32_t mip, uint32_t out[2]);
/* The pages the next frame uses (the page requests of VirtualCache::plan,
 * vcache.rs:3940-4050): a resident page is marked used; an absent one takes
 * a free slot, enters the page table and is planned for the next frame's
 * bake (every texel and light). slots[i]: its slot, or CNQ_NONE (no free
 * slot, a probe run past 64 entries, or a page outside its chart). */
int cnq_vcache_request(cnq_vcache* cache, const uint32_t* keys, uint32_t n, uint32_t* slots);
/* Frees resident pages (vcache.rs:3515 `evict_queued`'s removal). */
int cnq_vcache_evict(cnq_vcache* cache, const uint32_t* keys, uint32_t n);
/* Objects moved (the scene already moved: cnq_scene_apply_moves) for the
 * next frame's stale passes (VirtualCache::set_moved, vcache.rs:2426).
 * `scene_box` (6 floats) holds every texel's apex; NULL takes the object
 * hierarchy's root with the old boxes, padded (instanced.rs:1395-1406). */
int cnq_vcache_set_moved(cnq_vcache* cache, const cnq_moved* moved, uint32_t n, const float* scene_box);
/* One frame on `stream`: page-table updates (`vc_update`), the stale passes
 * if anything moved (`vc_stale`, `vc_stale_texels`), the bake of planned
 * and appended texels (`vc_bake`), the fill (`vc_fill`); then waits and
 * evicts the pages marked invalid (vcache.rs:3425). stats: [planned pages,
 * planned jobs, appended pages, appended jobs, candidates, invalid marks,
 * pages made stale for a re-bake, appends that did not fit, visible pages
 * tested, pages evicted]. */
int cnq_vcache_frame(cnq_vcache* cache, void* stream, uint64_t stats[10]);
/* Pages by key, as the shading read finds them (WGSL :3134 `vc_find`):
 * texels (81 x (lights + 1) floats per key, texel-major) and slots[i]
 * (CNQ_NONE, texels zero, where absent or marked stale). */
int cnq_vcache_read(cnq_vcache* cache, const uint32_t* keys, uint32_t n, float* texels, uint32_t* slots);
/* Shading from the cache (WGSL :5937 `vc_shade`, the direct light's part).
 * `hits` (device, 4 u32 per pixel as cnq_primary_device writes them) are a
 * camera pass of `camera`; per pixel its surface point and chart
 * coordinates pick the wanted mip (`vc_mip`: texels spanning
 * `pixels_per_texel` pixels of angle `pixel_angle`), then the page shown
 * (`vc_find_shown`: that mip's, else the finest coarser one resident), and
 * per light `out[l * w * h + p]` (device) is the page's share times the
 * unshadowed light at the pixel (or, without light_ratio, the cached value;
 * the unshadowed light where no page is resident). `out_page[p]`: the
 * shown page's mip (top bit: none found; CNQ_NONE: no surface). One pixel per
 * `request_tile` x `request_tile` (a different one each frame) requests the
 * page one to `refine_levels` mips finer than the one it found, or the
 * coarsest missing one (`vc_request`). counts: [requests, pixels without a
 * page]. */
typedef struct {
  float pixels_per_texel;  /* CacheOptions::pixels_per_texel (2) */
  uint32_t refine_levels;  /* CacheOptions::refine_levels (3), 1..4 */
  uint32_t request_tile;   /* CacheOptions::request_tile (4) */
  float pixel_angle;       /* 2 tan(fov / 2) / height */
} cnq_vcache_shade_params;
int cnq_vcache_shade(cnq_vcache* cache, const float* camera, uint32_t width, uint32_t height, const uint32_t* hits,
                     float* out, uint32_t* out_page, const cnq_vcache_shade_params* params, void* stream,
                     uint32_t counts[2]);
/* VirtualCache::plan (vcache.rs:3704-3810, 3928-4050) over the last shading
 * pass's requests: the non-resident pages by (priority, key), at most
 * `budget` cones of bake work (a page costs its texels inside the grid times
 * the lights; the first page always fits), pages unused for more than two
 * frames evicted least recently used first when no slot is free. The
 * planned pages bake in the next cnq_vcache_frame. stats: [requests, pages
 * planned, cones planned, pages evicted]. */
int cnq_vcache_plan(cnq_vcache* cache, uint32_t budget, uint64_t stats[4]);
/* A light edit of static lights (`Renderer::set_light`, renderer.rs:421,
 * for a light that is not moving): lights [first, first + n) replaced by
 * `lights` (n x 16 floats) and those lights' channels of every resident page
 * baked afresh, inside texels then border texels, so each one equals a
 * fresh bake. Refuses a moving light (cnq_vcache_move_light moves those).
 * stats (nullable): [pages, cones]. Waits for `stream`. */
int cnq_vcache_set_lights(cnq_vcache* cache, const float* lights, uint32_t first, uint32_t n, void* stream,
                          uint64_t stats[2]);

/* Moving lights through the error ledger (WGSL :4457-5164; vcache.rs
 * `VirtualCache::new`'s `moving` list, :1399-1430). Lights `lights[0..n)`
 * of the table move: each texel keeps its light's shadowed share of the
 * unshadowed light, and shading multiplies that by the unshadowed light
 * where the light is now. Each page keeps a 48-word record per moving
 * light: the frame it was baked at (its version), its slack (the smallest
 * guard cone's margin of its clear and facing-away texels, less the light's
 * shift since), and each texel's state (clear: nothing meets its guard cone,
 * three times the light's angular radius; mixed; or facing away). A rect
 * light bakes on a lattice fixed in its pose at the page's version, so a
 * mixed texel's ratio moves only by a bounded amount (`lattice_bound`). Each
 * frame the ledger pass re-bakes, for pages read in the frame before, the
 * mixed texels whose bound passes `epsilon`, or the whole page past a hashed
 * share (a half to all) of its slack; a page not read is set aside (kept,
 * shown as missing) until a pixel wants it. Ratios within `epsilon` of 0 or
 * 1 are stored as 0 or 1. Set once, before the first frame, with
 * light_ratio on; not together with moving objects (vcache.rs:1410). */
int cnq_vcache_set_moving_lights(cnq_vcache* cache, const uint32_t* lights, uint32_t n, float epsilon);
/* Moving light `index` is `light` (16 floats, the same kind: rect or
 * sphere) from the next frame (vcache.rs:2375 `set_light_pose`). */
int cnq_vcache_move_light(cnq_vcache* cache, uint32_t index, const float* light);
/* The last frame's ledger counters (WGSL mover_stats 8-14 and 21-23):
 * [pages re-baked, texels re-baked, pages set aside, pages checked (read
 * in the frame before), pages re-baking mixed texels, pages held whole,
 * jobs asked, pages held on their lattice, lattices that may have grown,
 * unbounded texels of held pages]. */
int cnq_vcache_ledger_stats(cnq_vcache* cache, uint64_t out[10]);
/* Diagnostics (either pointer may be NULL): the ledger records (pages x
 * moving lights x 48 words) and each record's certified bound (pages x
 * moving lights). */
int cnq_vcache_ledger_state(cnq_vcache* cache, uint32_t* records, float* bounds);
/* `verify_ledger` (instanced.rs:3741-3850) for moving light `light`: every
 * page shown baked afresh for that light (its records untouched) and
 * compared with the cache. out: [pages, pages whose largest change passes
 * their certified bound + 1e-5, pages past epsilon, pages changed at all,
 * the largest change, the largest bound]. */
int cnq_vcache_verify_ledger(cnq_vcache* cache, uint32_t light, void* stream, float out[6]);
/* `verify_cache` (instanced.rs:3721-3729): every page shown (resident, not
 * marked stale or invalid) baked afresh into a copy of the pool and compared
 * with the pool by f32 bit pattern. out: [values that differ, values
 * compared]. Refuses to run while planned pages wait for their bake. */
int cnq_vcache_verify(cnq_vcache* cache, void* stream, uint64_t out[2]);
/* Diagnostics (any pointer may be NULL): slot info (12 per slot), stale
 * marks, last use, the pool (81 x channels per slot), the device and host
 * page tables (4 per entry); sizes: [pages, channels, table entries, slots
 * ever used, next frame]. */
int cnq_vcache_state(cnq_vcache* cache, uint32_t* slot_info, uint32_t* slot_stale, uint32_t* slot_used, float* pool,
                     uint32_t* table_device, uint32_t* table_host, uint32_t* sizes);

#ifdef __cplusplus
}
#endif
#endif
