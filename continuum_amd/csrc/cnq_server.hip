// The served form: a request loop over one resident scene with continuous
// batching. Clients submit view requests (camera, size, light indices) and
// light edits from any thread. A light edit applies to the table at once; a
// view takes its lights from the table as it stands at submission. One
// scheduler thread admits views as HBM and launch slots free up, in arrival
// order, except that an express slot (the last, when configured) takes the
// small views first, packs the admitted views of each round into one camera
// pass and cone launches on a free slot's stream, and lets launches on
// different slots overlap on the device. A launch's cone work can be cut
// into pieces of a bounded number of queries, so the device turns to other
// streams' launches between pieces. A host callback on each slot's stream
// stamps completion, so a request's latency runs from submit to results
// readable.
//
// Memory is reserved once: the shading points and hits of admitted views
// live in a pixel arena (48 B per pixel), their query results in a query
// arena (20 B per query), each handed out in contiguous ranges by a
// first-fit allocator; admission is by these footprints, and a view's
// ranges are held until the client releases it. Nothing is allocated on
// the serving path.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>

#include "cnq_internal.h"

using namespace cnq::detail;
using cnq::NONE;

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

// First-fit ranges of [0, capacity), coalescing on release.
class RangeAlloc {
 public:
  void reset(uint64_t capacity) {
    free_.clear();
    if (capacity) free_[0] = capacity;
    capacity_ = capacity;
    used_ = 0;
  }
  bool take(uint64_t n, uint64_t& off) {
    if (n == 0) {
      off = 0;
      return true;
    }
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      off = it->first;
      const uint64_t rest = it->second - n;
      free_.erase(it);
      if (rest) free_[off + n] = rest;
      used_ += n;
      return true;
    }
    return false;
  }
  void give(uint64_t off, uint64_t n) {
    if (n == 0) return;
    used_ -= n;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + n == next->first) {
      n += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        prev->second += n;
        return;
      }
    }
    free_[off] = n;
  }
  uint64_t used() const { return used_; }

 private:
  std::map<uint64_t, uint64_t> free_;
  uint64_t capacity_ = 0, used_ = 0;
};

struct Request {
  uint64_t id = 0;
  bool is_view = true;
  float camera[20];
  uint32_t w = 0, h = 0;
  std::vector<uint32_t> lights;
  std::vector<float> rows;  // a view's lights, 16 floats each, as of submission
  int32_t state = CNQ_REQ_QUEUED;
  uint64_t px_off = 0, q_off = 0, launch = 0;
  Clock::time_point t_submit, t_admit, t_done;
  uint64_t pixels() const { return (uint64_t)w * h; }
  uint64_t queries() const { return pixels() * lights.size(); }
};

}  // namespace

struct cnq_server;

namespace {

struct Slot {
  cnq_server* owner = nullptr;
  uint32_t index = 0;
  hipStream_t stream = nullptr;
  DevBuf<cnq::ViewDesc> views;
  DevBuf<float> cameras, lights;
  DevBuf<uint32_t> work_counter;
  cnq::ViewDesc* h_views = nullptr;  // pinned staging
  float* h_cameras = nullptr;
  float* h_lights = nullptr;
  std::vector<Request*> reqs;  // the launch in flight
  uint64_t launch = 0;
  bool busy = false;
  bool finished = false;
  Clock::time_point t_done;
};

}  // namespace

struct cnq_server {
  cnq_scene* scene = nullptr;
  cnq_server_config cfg{};
  cnq_options opts{};
  std::vector<float> light_table;  // 16 floats per light
  DevBuf<float> receivers, visible;
  DevBuf<uint32_t> hits;
  DevBuf<uint64_t> full, touched;
  RangeAlloc px_alloc, q_alloc;
  std::vector<Slot> slots;

  std::mutex mu;
  std::condition_variable wake;  // the scheduler: new requests, finished launches, stop
  std::condition_variable done;  // clients: a request finished
  std::deque<Request*> pending;
  std::unordered_map<uint64_t, Request*> requests;
  uint64_t next_id = 1, launches = 0, views_done = 0, edits_done = 0, queries_done = 0;
  bool stop = false;
  Clock::time_point t_start = Clock::now();
  int error = CNQ_OK;
  std::string error_msg;
  std::thread worker;

  void release_buffers() {
    for (Slot& s : slots) {
      if (s.stream) (void)hipStreamDestroy(s.stream);
      s.stream = nullptr;
      s.views.release(), s.cameras.release(), s.lights.release(), s.work_counter.release();
      if (s.h_views) (void)hipHostFree(s.h_views);
      if (s.h_cameras) (void)hipHostFree(s.h_cameras);
      if (s.h_lights) (void)hipHostFree(s.h_lights);
      s.h_views = nullptr, s.h_cameras = nullptr, s.h_lights = nullptr;
    }
    receivers.release(), visible.release(), hits.release(), full.release(), touched.release();
  }
};

namespace {

// Stream callback: the slot's launch has finished on the device.
void slot_finished(void* arg) {
  Slot* s = static_cast<Slot*>(arg);
  const Clock::time_point now = Clock::now();
  std::lock_guard<std::mutex> g(s->owner->mu);
  s->t_done = now;
  s->finished = true;
  s->owner->wake.notify_all();
}

// Enqueues one launch of the views in `s.reqs` (their ranges assigned).
// Called without the lock.
hipError_t launch_slot(cnq_server* sv, Slot& s) {
  // View i's lights were copied to staging rows [i * max_view_lights, ...)
  // at admission, as the light table held them then.
  const uint32_t n = (uint32_t)s.reqs.size();
  const uint32_t tile_w = sv->cfg.tuning.tile_w, per_view = sv->cfg.max_view_lights;
  uint64_t ptiles = 0, items = 0;
  for (uint32_t i = 0; i < n; i++) {
    const Request& r = *s.reqs[i];
    cnq::ViewDesc& v = s.h_views[i];
    v = cnq::ViewDesc{};
    v.recv_off = r.px_off;
    v.out_off = r.q_off;
    v.w = r.w;
    v.h = r.h;
    v.light_off = i * per_view;
    v.n_lights = (uint32_t)r.lights.size();
    v.ptile_off = (uint32_t)ptiles;
    items += layout_view(v, tile_w, items);
    ptiles += (uint64_t)((r.w + 7) / 8) * ((r.h + 7) / 8);
    std::memcpy(s.h_cameras + 20 * (size_t)i, r.camera, 80);
  }
  if (ptiles > 0x7fffffffull || items > 0xffffffffull) return hipErrorInvalidValue;
  hipError_t e = hipMemcpyAsync(s.views.p, s.h_views, n * sizeof(cnq::ViewDesc), hipMemcpyHostToDevice, s.stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(s.cameras.p, s.h_cameras, 80 * (size_t)n, hipMemcpyHostToDevice, s.stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(s.lights.p, s.h_lights, 64 * (size_t)n * per_view, hipMemcpyHostToDevice, s.stream);
  if (e != hipSuccess) return e;
  cnq_scene* sc = sv->scene;
  cnq::PrimaryArgs pa;
  pa.s = sc->args();
  pa.cameras = s.cameras.p;
  pa.lights = s.lights.p;
  pa.views = s.views.p;
  pa.v0 = s.h_views[0];
  pa.n_views = n;
  pa.n_tiles = (uint32_t)ptiles;
  pa.hierarchy = sv->opts.hierarchy;
  pa.nearest_first = sv->opts.nearest_first;
  pa.hits = sv->hits.p;
  pa.receivers = sv->receivers.p;
  pa.counters = sc->counters.p + 3;
  pa.normals = sc->normals.p;
  e = cnq::launch_primary(pa, s.stream, sv->cfg.tuning.ray_stack ? sv->cfg.tuning.ray_stack : 64);
  if (e != hipSuccess) return e;
  if (items) {
    cnq::ItemArgs a = item_args(sc, sv->receivers.p, s.lights.p, sv->visible.p, sv->full.p, sv->touched.p, nullptr);
    a.query_flags = nullptr;  // the server's arena layout is its own
    a.n_flags = 0;
    a.p = query_params(sc, &sv->opts, tile_w);
    a.p.n_views = n;
    a.views = s.views.p;
    a.v0 = s.h_views[0];
    a.work_counter = s.work_counter.p;
    // Pieces of at most launch_queries queries (whole 64-query items), one
    // after another on this stream; a direct launch also stays within the
    // grid's 2^31 - 1 waves.
    uint64_t piece = sv->cfg.launch_queries ? (sv->cfg.launch_queries + 63) / 64 : items;
    if (!sv->cfg.tuning.persistent_blocks && piece > 0x7fffffffull) piece = 0x7fffffffull;
    for (uint64_t b = 0; b < items; b += piece) {
      a.p.item_begin = (uint32_t)b;
      a.p.n_items = (uint32_t)(items - b < piece ? items : b + piece);
      if (sv->cfg.tuning.persistent_blocks) {
        e = fill_u32(s.work_counter.p, 0, 1, s.stream);
        if (e != hipSuccess) return e;
      }
      e = cnq::launch_cone_items(a, sc->count_work, sv->cfg.tuning.persistent_blocks, s.stream,
                                 sv->cfg.tuning.waves_per_simd);
      if (e != hipSuccess) return e;
    }
  }
  return hipLaunchHostFunc(s.stream, slot_finished, &s);
}

void scheduler(cnq_server* sv) {
  (void)hipSetDevice(sv->scene->dev->ordinal);
  std::unique_lock<std::mutex> lk(sv->mu);
  for (;;) {
    // Retire finished launches.
    for (Slot& s : sv->slots) {
      if (!s.busy || !s.finished) continue;
      for (Request* r : s.reqs) {
        r->state = CNQ_REQ_DONE;
        r->t_done = s.t_done;
        sv->views_done++;
        sv->queries_done += r->queries();
      }
      s.reqs.clear();
      s.busy = false;
      s.finished = false;
      sv->done.notify_all();
    }
    bool any_busy = false;
    for (Slot& s : sv->slots) any_busy = any_busy || s.busy;
    if (sv->stop && !any_busy) break;
    // Admit views into free slots, in arrival order; the express slot
    // passes over views larger than express_queries. A view that does not
    // fit the arenas ends the slot's round.
    bool launched = false;
    if (!sv->stop && sv->error == CNQ_OK) {
      for (size_t si = 0; si < sv->slots.size(); si++) {
        Slot& s = sv->slots[si];
        if (s.busy || sv->pending.empty()) continue;
        const bool express = sv->cfg.express_queries && si + 1 == sv->slots.size();
        for (auto it = sv->pending.begin(); it != sv->pending.end();) {
          Request* r = *it;
          if (s.reqs.size() >= sv->cfg.max_views_per_launch) break;
          if (express && r->queries() > sv->cfg.express_queries) {
            ++it;
            continue;
          }
          uint64_t po, qo;
          if (!sv->px_alloc.take(r->pixels(), po)) break;
          if (!sv->q_alloc.take(r->queries(), qo)) {
            sv->px_alloc.give(po, r->pixels());
            break;
          }
          r->px_off = po;
          r->q_off = qo;
          if (!r->rows.empty())
            std::memcpy(s.h_lights + 16 * s.reqs.size() * sv->cfg.max_view_lights, r->rows.data(),
                        4 * r->rows.size());
          r->state = CNQ_REQ_ADMITTED;
          r->t_admit = Clock::now();
          r->launch = sv->launches + 1;
          s.reqs.push_back(r);
          it = sv->pending.erase(it);
        }
        if (s.reqs.empty()) continue;
        s.launch = ++sv->launches;
        s.busy = true;
        s.finished = false;
        lk.unlock();
        const hipError_t e = launch_slot(sv, s);
        lk.lock();
        if (e != hipSuccess) {
          // The launch never ran: report the failure on its requests.
          sv->error = CNQ_HIP_ERROR;
          sv->error_msg = std::string("server launch: ") + hipGetErrorString(e);
          s.busy = false;
          s.reqs.clear();
          sv->done.notify_all();
        }
        launched = true;
      }
    }
    if (launched) continue;
    sv->wake.wait(lk);
  }
}

}  // namespace

extern "C" {

void cnq_server_config_default(cnq_server_config* c) {
  c->pixel_capacity = 16ull << 20;   // 768 MiB of shading points and hits
  c->query_capacity = 128ull << 20;  // 2.5 GiB of results
  c->max_lights = 256;
  c->max_view_lights = 32;
  c->max_views_per_launch = 64;
  c->slots = 2;
  cnq_tuning_default(&c->tuning);
  c->launch_queries = 0;
  c->express_queries = 0;
}

int cnq_server_create(cnq_scene* scene, const cnq_server_config* cfg, const cnq_options* opts, const float* lights,
                      uint32_t n_lights, cnq_server** out) {
  if (!scene || !cfg || !opts || !out) return fail(CNQ_INVALID_ARG, "null argument");
  int rc = check_options(opts);
  if (rc || (rc = check_object_cov(scene, opts))) return rc;
  if (cfg->slots == 0 || cfg->slots > 8) return fail(CNQ_INVALID_ARG, "slots must be 1..8");
  if (cfg->max_views_per_launch == 0 || cfg->max_view_lights == 0 || cfg->max_lights == 0)
    return fail(CNQ_INVALID_ARG, "zero-sized server limits");
  if (cfg->express_queries && cfg->slots < 2)
    return fail(CNQ_INVALID_ARG, "an express slot needs at least two slots");
  const cnq_tuning& t = cfg->tuning;
  if (t.tile_w == 0 || t.tile_w > 64 || (t.tile_w & (t.tile_w - 1)))
    return fail(CNQ_INVALID_ARG, "tile_w must be 1..64, a power of two");
  if (n_lights > cfg->max_lights || (n_lights && !lights)) return fail(CNQ_INVALID_ARG, "light table");
  if (cfg->pixel_capacity > 0x7fffffffull * 8 || cfg->query_capacity > (1ull << 34))
    return fail(CNQ_CAPACITY, "server capacity too large");
  HIP_TRY(hipSetDevice(scene->dev->ordinal), "hipSetDevice");
  cnq_server* sv = new (std::nothrow) cnq_server;
  if (!sv) return fail(CNQ_OOM, "out of host memory");
  sv->scene = scene;
  sv->cfg = *cfg;
  sv->opts = *opts;
  sv->light_table.assign(16 * (size_t)cfg->max_lights, 0.0f);
  if (n_lights) std::memcpy(sv->light_table.data(), lights, 64 * (size_t)n_lights);
  sv->px_alloc.reset(cfg->pixel_capacity);
  sv->q_alloc.reset(cfg->query_capacity);
  hipError_t e = sv->receivers.alloc(8 * cfg->pixel_capacity);
  if (e == hipSuccess) e = sv->hits.alloc(4 * cfg->pixel_capacity);
  if (e == hipSuccess) e = sv->visible.alloc(cfg->query_capacity);
  if (e == hipSuccess) e = sv->full.alloc(cfg->query_capacity);
  if (e == hipSuccess) e = sv->touched.alloc(cfg->query_capacity);
  sv->slots.resize(cfg->slots);
  const size_t nv = cfg->max_views_per_launch, nl = nv * cfg->max_view_lights;
  for (uint32_t i = 0; i < cfg->slots && e == hipSuccess; i++) {
    Slot& s = sv->slots[i];
    s.owner = sv;
    s.index = i;
    e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = s.views.alloc(nv);
    if (e == hipSuccess) e = s.cameras.alloc(20 * nv);
    if (e == hipSuccess) e = s.lights.alloc(16 * nl);
    if (e == hipSuccess) e = s.work_counter.alloc(1);
    if (e == hipSuccess) e = hipHostMalloc((void**)&s.h_views, nv * sizeof(cnq::ViewDesc), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&s.h_cameras, 80 * nv, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&s.h_lights, 64 * nl, hipHostMallocDefault);
  }
  if (e != hipSuccess) {
    sv->release_buffers();
    delete sv;
    return hip_fail(e, e == hipErrorOutOfMemory ? "server buffers (out of device memory)" : "server buffers");
  }
  try {
    sv->worker = std::thread(scheduler, sv);
  } catch (...) {
    sv->release_buffers();
    delete sv;
    return fail(CNQ_OOM, "could not start the scheduler thread");
  }
  *out = sv;
  return CNQ_OK;
}

void cnq_server_destroy(cnq_server* sv) {
  if (!sv) return;
  {
    std::lock_guard<std::mutex> g(sv->mu);
    sv->stop = true;
    sv->wake.notify_all();
  }
  if (sv->worker.joinable()) sv->worker.join();
  (void)hipSetDevice(sv->scene->dev->ordinal);
  for (Slot& s : sv->slots)
    if (s.stream) (void)hipStreamSynchronize(s.stream);
  for (auto& kv : sv->requests) delete kv.second;
  sv->requests.clear();
  sv->release_buffers();
  delete sv;
}

int cnq_server_submit_view(cnq_server* sv, const float* camera, uint32_t width, uint32_t height,
                           const uint32_t* lights, uint32_t n_lights, uint64_t* id) {
  if (!sv || !camera || !id || (n_lights && !lights)) return fail(CNQ_INVALID_ARG, "null argument");
  if (width == 0 || height == 0) return fail(CNQ_INVALID_ARG, "empty view");
  if (n_lights > sv->cfg.max_view_lights) return fail(CNQ_CAPACITY, "too many lights for one view");
  for (uint32_t k = 0; k < n_lights; k++)
    if (lights[k] >= sv->cfg.max_lights) return fail(CNQ_INVALID_ARG, "light index outside the table");
  const uint64_t px = (uint64_t)width * height;
  if (px > sv->cfg.pixel_capacity || px * n_lights > sv->cfg.query_capacity)
    return fail(CNQ_CAPACITY, "view larger than the server's capacity");
  Request* r = new (std::nothrow) Request;
  if (!r) return fail(CNQ_OOM, "out of host memory");
  std::memcpy(r->camera, camera, 80);
  r->w = width;
  r->h = height;
  r->lights.assign(lights, lights + n_lights);
  r->rows.resize(16 * (size_t)n_lights);
  r->t_submit = Clock::now();
  std::lock_guard<std::mutex> g(sv->mu);
  if (sv->stop) {
    delete r;
    return fail(CNQ_INVALID_ARG, "server is stopping");
  }
  if (sv->error) {
    delete r;
    return fail(sv->error, sv->error_msg);
  }
  for (uint32_t k = 0; k < n_lights; k++)
    std::memcpy(r->rows.data() + 16 * (size_t)k, &sv->light_table[16 * (size_t)lights[k]], 64);
  r->id = sv->next_id++;
  sv->requests[r->id] = r;
  sv->pending.push_back(r);
  sv->wake.notify_all();
  *id = r->id;
  return CNQ_OK;
}

int cnq_server_submit_light(cnq_server* sv, uint32_t index, const float* light, uint64_t* id) {
  if (!sv || !light || !id) return fail(CNQ_INVALID_ARG, "null argument");
  if (index >= sv->cfg.max_lights) return fail(CNQ_INVALID_ARG, "light index outside the table");
  Request* r = new (std::nothrow) Request;
  if (!r) return fail(CNQ_OOM, "out of host memory");
  r->is_view = false;
  r->t_submit = Clock::now();
  std::lock_guard<std::mutex> g(sv->mu);
  if (sv->stop) {
    delete r;
    return fail(CNQ_INVALID_ARG, "server is stopping");
  }
  // Applied at once: views submitted from now on see the new light.
  std::memcpy(&sv->light_table[16 * (size_t)index], light, 64);
  r->state = CNQ_REQ_DONE;
  r->t_admit = r->t_done = Clock::now();
  sv->edits_done++;
  r->id = sv->next_id++;
  sv->requests[r->id] = r;
  sv->done.notify_all();
  *id = r->id;
  return CNQ_OK;
}

int cnq_server_wait(cnq_server* sv, uint64_t id, double timeout_ms, int* done) {
  if (!sv || !done) return fail(CNQ_INVALID_ARG, "null argument");
  std::unique_lock<std::mutex> lk(sv->mu);
  auto it = sv->requests.find(id);
  if (it == sv->requests.end()) return fail(CNQ_INVALID_ARG, "no request " + std::to_string(id));
  Request* r = it->second;
  const auto until = Clock::now() + std::chrono::duration_cast<Clock::duration>(
                                        std::chrono::duration<double, std::milli>(timeout_ms < 0 ? 0 : timeout_ms));
  while (r->state < CNQ_REQ_DONE && !sv->error && !sv->stop) {
    if (sv->done.wait_until(lk, until) == std::cv_status::timeout) break;
  }
  *done = r->state >= CNQ_REQ_DONE ? 1 : 0;
  if (!*done && sv->error) return fail(sv->error, sv->error_msg);
  return CNQ_OK;
}

int cnq_server_info(cnq_server* sv, uint64_t id, cnq_request_info* info) {
  if (!sv || !info) return fail(CNQ_INVALID_ARG, "null argument");
  std::lock_guard<std::mutex> g(sv->mu);
  auto it = sv->requests.find(id);
  if (it == sv->requests.end()) return fail(CNQ_INVALID_ARG, "no request " + std::to_string(id));
  const Request* r = it->second;
  info->state = r->state;
  info->width = r->w;
  info->height = r->h;
  info->n_lights = (uint32_t)r->lights.size();
  info->launch = r->launch;
  info->queued_ms = r->state >= CNQ_REQ_ADMITTED ? ms_between(r->t_submit, r->t_admit) : -1.0;
  info->latency_ms = r->state >= CNQ_REQ_DONE ? ms_between(r->t_submit, r->t_done) : -1.0;
  info->admitted_at_ms = r->state >= CNQ_REQ_ADMITTED ? ms_between(sv->t_start, r->t_admit) : -1.0;
  info->done_at_ms = r->state >= CNQ_REQ_DONE ? ms_between(sv->t_start, r->t_done) : -1.0;
  return CNQ_OK;
}

int cnq_server_read(cnq_server* sv, uint64_t id, uint32_t* hits, float* visible, uint64_t* full, uint64_t* touched) {
  if (!sv) return fail(CNQ_INVALID_ARG, "server is null");
  uint64_t po, qo, px, q;
  {
    std::lock_guard<std::mutex> g(sv->mu);
    auto it = sv->requests.find(id);
    if (it == sv->requests.end()) return fail(CNQ_INVALID_ARG, "no request " + std::to_string(id));
    const Request* r = it->second;
    if (!r->is_view) return fail(CNQ_INVALID_ARG, "a light edit has no results");
    if (r->state != CNQ_REQ_DONE) return fail(CNQ_INVALID_ARG, "request not done, or released");
    po = r->px_off, qo = r->q_off, px = r->pixels(), q = r->queries();
  }
  // The ranges stay held until release, so the copies need no lock.
  HIP_TRY(hipSetDevice(sv->scene->dev->ordinal), "hipSetDevice");
  if (hits) HIP_TRY(hipMemcpy(hits, sv->hits.p + 4 * po, 16 * px, hipMemcpyDeviceToHost), "read hits");
  if (q && visible) HIP_TRY(hipMemcpy(visible, sv->visible.p + qo, 4 * q, hipMemcpyDeviceToHost), "read visible");
  if (q && full) HIP_TRY(hipMemcpy(full, sv->full.p + qo, 8 * q, hipMemcpyDeviceToHost), "read full");
  if (q && touched) HIP_TRY(hipMemcpy(touched, sv->touched.p + qo, 8 * q, hipMemcpyDeviceToHost), "read touched");
  return CNQ_OK;
}

int cnq_server_release(cnq_server* sv, uint64_t id) {
  if (!sv) return fail(CNQ_INVALID_ARG, "server is null");
  std::lock_guard<std::mutex> g(sv->mu);
  auto it = sv->requests.find(id);
  if (it == sv->requests.end()) return fail(CNQ_INVALID_ARG, "no request " + std::to_string(id));
  Request* r = it->second;
  if (r->state != CNQ_REQ_DONE) return fail(CNQ_INVALID_ARG, "only a done request can be released");
  if (r->is_view) {
    sv->px_alloc.give(r->px_off, r->pixels());
    sv->q_alloc.give(r->q_off, r->queries());
  }
  sv->requests.erase(it);
  delete r;
  sv->wake.notify_all();
  return CNQ_OK;
}

int cnq_server_counts(cnq_server* sv, uint64_t counts[6]) {
  if (!sv || !counts) return fail(CNQ_INVALID_ARG, "null argument");
  std::lock_guard<std::mutex> g(sv->mu);
  counts[0] = sv->views_done;
  counts[1] = sv->edits_done;
  counts[2] = sv->launches;
  counts[3] = sv->queries_done;
  counts[4] = sv->px_alloc.used();
  counts[5] = sv->q_alloc.used();
  return CNQ_OK;
}

}  // extern "C"
