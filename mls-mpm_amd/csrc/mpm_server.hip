// mpm_server.hip -- serving form: a request loop with continuous batching.
//
// Requests (a scene configuration, its particles and a step count) are
// queued by any thread.  One loop thread owns the device: it admits queued
// requests in arrival order while the HBM budget and the resident-scene
// limit allow (creating and loading each scene while the device keeps
// stepping the others), keeps every admitted scene resident in one mpm_batch,
// advances all of them together in chunks of batched steps (one persistent
// launch per stage), and retires each request when its steps are done: the
// final particles are copied to host memory owned by the request and the
// scene's device memory is released, making room for the next admissions.
// A request's result is bit-identical to stepping its scene alone: the
// batched kernels run the single-scene bodies.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "mpm_scene.h"

using namespace mpm;

namespace {

using Clock = std::chrono::steady_clock;

// Drift margin of every served scene (the members of one batch share it).
constexpr uint32_t kServedMargin = 2;
// Arena slab: scenes larger than this take a slab of their own size.
constexpr size_t kArenaSlab = (size_t)1 << 30;
// Pinned result arena: slab size and cap (results beyond it go to pageable
// vectors until the clients fetch theirs).
constexpr size_t kResultSlab = (size_t)256 << 20;
constexpr size_t kResultCap = (size_t)16 << 30;

struct Request {
  uint64_t id = 0;
  mpm_config cfg{};
  uint32_t n = 0, steps = 0, done = 0;
  std::vector<float> pos3, vel3, C9, mass;
  int state = MPM_REQ_QUEUED;
  bool cancel = false;
  std::string error;
  size_t footprint = 0;
  mpm_scene *scene = nullptr;
  // The final particles: a pinned block of the server's result arena (pos4,
  // vel4, C9, density: 18 floats per particle), or, when that arena is full,
  // these vectors.
  float *res = nullptr;
  std::vector<float> pos4, vel4, C9out, density;
  double t_submit = 0, t_admit = 0, t_done = 0;
};

int sfail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int sfail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  return set_error(code, buf);
}

}  // namespace

struct mpm_server {
  mpm_server_config cfg{};
  Clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv_work;  // loop: new request, cancel or stop
  std::condition_variable cv_done;  // waiters: a request finished
  bool stop = false;
  uint64_t next_id = 1;
  std::deque<uint64_t> queue;                          // admission order
  std::map<uint64_t, std::unique_ptr<Request>> reqs;   // every unreleased request
  std::vector<Request *> active;                       // resident (changed by the loop, under mu)
  uint64_t bytes_resident = 0;
  mpm_server_stats stats{};
  std::vector<float> step_ms;                          // per batched step
  std::vector<mpm_step_record> step_rec;               // per batched step
  std::thread loop;
  mpm_batch *batch = nullptr;
  // device memory of the resident scenes: slabs kept for the server's life
  std::unique_ptr<Arena> arena;
  // pinned host memory the finished requests' particles are read into (one
  // fast copy on the loop thread), until their results are fetched
  std::unique_ptr<Arena> results;
  int loop_error = 0;
  std::string loop_msg;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

// Create and load one request's scene (its buffers from the server's arena;
// the load sorts the caller's order into bricks on the scene's own stream),
// while the batch runs (loop thread).  Every step then runs in the batch.
int prepare(mpm_server *srv, Request *r) {
  const double t0 = srv->now();
  int rc = validate_config(r->cfg);
  if (!rc) {
    r->scene = new mpm_scene();
    rc = scene_init(r->scene, r->cfg, srv->arena.get());
  }
  const double t1 = srv->now();
  if (!rc)
    rc = mpm_reset(r->scene, r->n, r->pos3.data(), r->vel3.empty() ? nullptr : r->vel3.data(),
                   r->C9.empty() ? nullptr : r->C9.data(),
                   r->mass.empty() ? nullptr : r->mass.data());
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stats.admit_create_seconds += t1 - t0;
    srv->stats.admit_load_seconds += srv->now() - t1;
  }
  if (rc) {
    std::string msg = mpm_last_error();
    if (r->scene) mpm_destroy(r->scene);
    r->scene = nullptr;
    std::lock_guard<std::mutex> lk(srv->mu);
    r->state = MPM_REQ_FAILED;
    r->error = msg;
    r->t_done = srv->now();
    srv->bytes_resident -= r->footprint;
    srv->stats.failed++;
    srv->cv_done.notify_all();
    return rc;
  }
  // inputs are on the device now
  std::vector<float>().swap(r->pos3);
  std::vector<float>().swap(r->vel3);
  std::vector<float>().swap(r->C9);
  std::vector<float>().swap(r->mass);
  return MPM_OK;
}

// Join a prepared scene to the batch (after the chunk it overlapped).
int enroll(mpm_server *srv, Request *r) {
  int rc = mpm_batch_add(srv->batch, r->scene);
  if (rc) {
    std::string msg = mpm_last_error();
    mpm_destroy(r->scene);
    r->scene = nullptr;
    std::lock_guard<std::mutex> lk(srv->mu);
    r->state = MPM_REQ_FAILED;
    r->error = msg;
    r->t_done = srv->now();
    srv->bytes_resident -= r->footprint;
    srv->stats.failed++;
    srv->cv_done.notify_all();
    return rc;
  }
  std::lock_guard<std::mutex> lk(srv->mu);
  srv->active.push_back(r);
  return MPM_OK;
}

// Read the final state and release the scene (loop thread).  The scene has
// left the batch: its read runs on its own stream, beside the next chunk,
// and its buffers go back to the arena.
void retire(mpm_server *srv, Request *r, bool cancelled, hipEvent_t copied = nullptr) {
  const double t0 = srv->now();
  int rc = MPM_OK;
  std::string msg;
  if (copied) {
    // read back already: wait for its copy
    const hipError_t e = hipEventSynchronize(copied);
    if (e != hipSuccess) {
      rc = hip_fail(e, "result copy");
      msg = mpm_last_error();
      srv->results->release(r->res);
      r->res = nullptr;
    }
  } else if (!cancelled) {
    const size_t n = r->n;
    void *blk = nullptr;
    if (n && srv->results->alloc(n * 18 * sizeof(float), &blk) == 0) {
      r->res = (float *)blk;
      rc = mpm_read_particles(r->scene, r->res, r->res + 4 * n, r->res + 8 * n, r->res + 17 * n);
      if (rc) {
        srv->results->release(blk);
        r->res = nullptr;
      }
    } else {
      r->pos4.resize(n * 4);
      r->vel4.resize(n * 4);
      r->C9out.resize(n * 9);
      r->density.resize(n);
      rc = mpm_read_particles(r->scene, r->pos4.data(), r->vel4.data(), r->C9out.data(),
                              r->density.data());
    }
    if (rc) msg = mpm_last_error();
  }
  const double t1 = srv->now();
  mpm_destroy(r->scene);
  r->scene = nullptr;
  std::lock_guard<std::mutex> lk(srv->mu);
  srv->stats.retire_read_seconds += t1 - t0;
  srv->stats.retire_release_seconds += srv->now() - t1;
  srv->bytes_resident -= r->footprint;
  r->t_done = srv->now();
  if (cancelled) {
    r->state = MPM_REQ_CANCELLED;
    srv->stats.cancelled++;
  } else if (rc) {
    r->state = MPM_REQ_FAILED;
    r->error = msg;
    srv->stats.failed++;
  } else {
    r->state = MPM_REQ_DONE;
    srv->stats.completed++;
  }
  srv->cv_done.notify_all();
}

// A request that left the batch after a chunk.  Its particles are exported
// on the batch's stream before the next chunk and copied into its pinned
// result block on the copy stream beside it (`issued`, events ev[0]: the
// export, ev[1]: the copy); it is released once the copy is done.
struct Leaving {
  Request *r;
  bool cancelled;
  bool issued = false;
  hipEvent_t ev[2] = {nullptr, nullptr};
};

void run_loop(mpm_server *srv) {
  if (hipSetDevice(srv->cfg.device) != hipSuccess || mpm_batch_create(srv->cfg.device, &srv->batch)) {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->loop_error = MPM_ERR_HIP;
    srv->loop_msg = mpm_last_error();
    srv->stop = true;
    srv->cv_done.notify_all();
    return;
  }
  mpm_batch_tuning tune{};
  mpm_batch_get_tuning(srv->batch, &tune);
  tune.item_blocks = srv->cfg.item_blocks;
  if (srv->cfg.resolve_blocks) tune.resolve_blocks = srv->cfg.resolve_blocks;
  if (srv->cfg.aux_blocks) tune.aux_blocks = srv->cfg.aux_blocks;
  tune.flags = srv->cfg.batch_flags ? srv->cfg.batch_flags : MPM_BATCH_GRAPH_REUSED;
  mpm_batch_set_tuning(srv->batch, &tune);
  mpm_batch_enable_timing(srv->batch, 1);
  const uint32_t chunk = srv->cfg.chunk_steps ? srv->cfg.chunk_steps : 8;
  std::vector<float> ms;
  // Finished requests' particles go to the host on their own stream, beside
  // the batch's next chunk.
  hipStream_t copy = nullptr;
  if (hipStreamCreateWithFlags(&copy, hipStreamNonBlocking) != hipSuccess) copy = nullptr;
  std::vector<hipEvent_t> evpool;
  auto take_event = [&]() -> hipEvent_t {
    hipEvent_t e = nullptr;
    if (!evpool.empty()) {
      e = evpool.back();
      evpool.pop_back();
    } else if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      e = nullptr;
    }
    return e;
  };
  // requests that left the batch after the last chunk, retired while the
  // next one runs
  std::vector<Leaving> leaving;
  // Export and copy of a leaving request's particles, enqueued before the
  // next chunk (the batch's stream is idle: the export runs at once, the
  // copy beside the chunk); left to retire's own read when it cannot be.
  auto issue_read = [&](Leaving &l) {
    const size_t n = l.r->n;
    if (l.cancelled || !n || !copy) return;
    void *blk = nullptr;
    if (srv->results->alloc(n * 18 * sizeof(float), &blk) != 0) return;
    l.ev[0] = take_event();
    l.ev[1] = take_event();
    if (l.ev[0] && l.ev[1] &&
        scene_export_async(l.r->scene, batch_stream(srv->batch), copy, blk, l.ev[0], l.ev[1]) ==
            MPM_OK) {
      l.r->res = (float *)blk;
      l.issued = true;
    } else {
      srv->results->release(blk);
    }
  };
  auto retire_leaving = [&]() {
    const double t0 = srv->now();
    for (Leaving &l : leaving) {
      retire(srv, l.r, l.cancelled, l.issued ? l.ev[1] : nullptr);
      for (hipEvent_t e : l.ev)
        if (e) evpool.push_back(e);
    }
    const bool any = !leaving.empty();
    leaving.clear();
    if (any) {
      std::lock_guard<std::mutex> lk(srv->mu);
      srv->stats.retire_seconds += srv->now() - t0;
    }
  };
  for (;;) {
    // Admissions, in arrival order, while the budget and slot limit allow.
    std::vector<Request *> admitting;
    {
      std::unique_lock<std::mutex> lk(srv->mu);
      if (srv->active.empty() && srv->queue.empty() && leaving.empty()) {
        if (srv->stop) break;
        srv->cv_work.wait_for(lk, std::chrono::milliseconds(20));
        continue;
      }
      while (!srv->queue.empty()) {
        Request *r = srv->reqs[srv->queue.front()].get();
        if (r->cancel) {
          srv->queue.pop_front();
          r->state = MPM_REQ_CANCELLED;
          r->t_done = srv->now();
          srv->stats.cancelled++;
          srv->cv_done.notify_all();
          continue;
        }
        const size_t resident = srv->active.size() + admitting.size();
        if (srv->cfg.max_active && resident >= srv->cfg.max_active) break;
        if (srv->bytes_resident + r->footprint > srv->cfg.hbm_budget) break;
        srv->queue.pop_front();
        srv->bytes_resident += r->footprint;
        r->state = MPM_REQ_RUNNING;
        r->t_admit = srv->now();
        admitting.push_back(r);
      }
    }
    // One chunk of batched steps over every resident scene, enqueued first;
    // the requests that finished in the last chunk are read back and
    // released, and the admitted scenes created, loaded and sorted, while
    // it runs.
    const bool stepping = !srv->active.empty();
    uint32_t c = chunk;
    for (Request *r : srv->active) c = std::min(c, r->steps - r->done);
    const double tb = srv->now();
    int rc = stepping ? mpm_batch_step(srv->batch, c) : MPM_OK;
    const double t_enq = srv->now();
    retire_leaving();
    const double t_pre = srv->now();
    std::vector<Request *> prepared;
    for (Request *r : admitting)
      if (prepare(srv, r) == MPM_OK) prepared.push_back(r);
    const double t_prep = srv->now();
    auto join_prepared = [&]() {
      const double t0 = srv->now();
      for (Request *r : prepared) {
        if (r->done >= r->steps) retire(srv, r, false);
        else enroll(srv, r);
      }
      std::lock_guard<std::mutex> lk(srv->mu);
      srv->stats.enroll_seconds += srv->now() - t0;
    };
    {
      std::lock_guard<std::mutex> lk(srv->mu);
      srv->stats.enqueue_seconds += t_enq - tb;
      srv->stats.admit_seconds += t_prep - t_pre;
      srv->stats.admitted += admitting.size();
    }
    if (!stepping) {
      join_prepared();
      continue;
    }
    if (!rc) rc = mpm_batch_sync(srv->batch);
    const double t_sync = srv->now();
    uint32_t cnt = 0;
    if (!rc) rc = mpm_batch_step_times(srv->batch, nullptr, 0, &cnt);
    if (!rc && cnt) {
      ms.resize(cnt);
      rc = mpm_batch_step_times(srv->batch, ms.data(), cnt, &cnt);
      mpm_batch_enable_timing(srv->batch, 1);  // restart the record
    }
    {
      std::lock_guard<std::mutex> lk(srv->mu);
      srv->stats.wait_seconds += t_sync - t_prep;
      srv->stats.chunks++;
      for (uint32_t i = 0; i < cnt && i < ms.size(); i++) srv->stats.gpu_step_seconds += ms[i] * 1e-3;
      batch_graph_counts(srv->batch, &srv->stats.graph_steps, &srv->stats.graph_captures);
    }
    if (rc) {
      std::string msg = mpm_last_error();
      std::vector<Request *> act;
      {
        std::lock_guard<std::mutex> lk(srv->mu);
        act.swap(srv->active);
      }
      for (Request *r : act) {
        (void)mpm_batch_remove(srv->batch, r->scene);
        mpm_destroy(r->scene);
        r->scene = nullptr;
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->bytes_resident -= r->footprint;
        r->state = MPM_REQ_FAILED;
        r->error = msg;
        r->t_done = srv->now();
        srv->stats.failed++;
      }
      {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->cv_done.notify_all();
      }
      join_prepared();
      continue;
    }
    std::vector<Request *> rest;
    {
      std::lock_guard<std::mutex> lk(srv->mu);
      uint64_t ps = 0, np = 0;
      for (Request *r : srv->active) {
        np += r->n;
        r->done += c;
        ps += (uint64_t)r->n * c;
        if (r->cancel) leaving.push_back({r, true});
        else if (r->done >= r->steps) leaving.push_back({r, false});
        else rest.push_back(r);
      }
      srv->stats.particle_steps += ps;
      srv->stats.batched_steps += c;
      srv->stats.resident_steps += (uint64_t)srv->active.size() * c;
      srv->stats.busy_seconds += t_sync - tb;
      srv->step_ms.insert(srv->step_ms.end(), ms.begin(), ms.begin() + std::min<size_t>(cnt, ms.size()));
      for (uint32_t i = 0; i < cnt && i < ms.size(); i++) {
        mpm_step_record q{};
        q.ms = ms[i];
        q.resident = (uint32_t)srv->active.size();
        q.admitting = (uint32_t)admitting.size();
        q.resident_particles = np;
        srv->step_rec.push_back(q);
      }
    }
    // the finished and cancelled leave the batch now (its stream is idle),
    // their particles are exported ahead of the next chunk, and they are
    // retired while it runs
    for (Leaving &l : leaving) {
      (void)mpm_batch_remove(srv->batch, l.r->scene);
      issue_read(l);
    }
    {
      std::lock_guard<std::mutex> lk(srv->mu);
      srv->active.swap(rest);
    }
    join_prepared();
  }
  retire_leaving();
  for (hipEvent_t e : evpool) (void)hipEventDestroy(e);
  if (copy) (void)hipStreamDestroy(copy);
  mpm_batch_destroy(srv->batch);
  srv->batch = nullptr;
}

}  // namespace

extern "C" {

int mpm_server_create(const mpm_server_config *cfg, mpm_server **out) {
  if (!cfg || !out) return sfail(MPM_ERR_INVALID, "mpm_server_create: NULL argument");
  *out = nullptr;
  if (!cfg->hbm_budget) return sfail(MPM_ERR_INVALID, "mpm_server_create: hbm_budget is 0");
  for (int i = 0; i < 16; i++)
    if (cfg->tile_ppt_by_class[i] && !tiled_ppt_supported(cfg->tile_ppt_by_class[i]))
      return sfail(MPM_ERR_INVALID, "mpm_server_create: tile_ppt_by_class[%d] = %u (4, 8 or 16)",
                   i, (unsigned)cfg->tile_ppt_by_class[i]);
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0)
    return sfail(MPM_ERR_HIP, "no HIP device available (%s)",
                 e == hipSuccess ? "0 devices" : hipGetErrorString(e));
  if (cfg->device < 0 || cfg->device >= ndev)
    return sfail(MPM_ERR_INVALID, "device %d out of range (%d devices)", cfg->device, ndev);
  mpm_server *srv = new mpm_server();
  srv->cfg = *cfg;
  srv->arena = std::make_unique<Arena>(cfg->device, kArenaSlab);
  srv->results = std::make_unique<Arena>(cfg->device, kResultSlab, true, kResultCap);
  srv->t0 = Clock::now();
  srv->loop = std::thread(run_loop, srv);
  *out = srv;
  return MPM_OK;
}

int mpm_server_destroy(mpm_server *srv) {
  if (!srv) return MPM_OK;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stop = true;
    for (auto &kv : srv->reqs) kv.second->cancel = true;
    srv->cv_work.notify_all();
  }
  if (srv->loop.joinable()) srv->loop.join();
  delete srv;
  return MPM_OK;
}

int mpm_server_submit(mpm_server *srv, const mpm_config *cfg, uint32_t n, const float *pos3,
                      const float *vel3, const float *C9, const float *mass, uint32_t steps,
                      uint64_t *id) {
  if (!srv || !cfg || !id || (n && !pos3))
    return sfail(MPM_ERR_INVALID, "mpm_server_submit: NULL argument");
  mpm_config c = *cfg;
  c.device = srv->cfg.device;
  c.sort_period = 1;
  c.flags &= ~(MPM_FLAG_NO_LDS | MPM_FLAG_KEEP_GRID);
  if (c.diffuse_capacity)
    return sfail(MPM_ERR_UNSUPPORTED, "served scenes carry no diffuse pool");
  // every resident scene shares one batch, whose members share one drift
  // margin: served scenes all take the default
  if (c.tile_margin && c.tile_margin != kServedMargin)
    return sfail(MPM_ERR_UNSUPPORTED, "served scenes use tile_margin %u (got %u)", kServedMargin,
                 c.tile_margin);
  c.tile_margin = kServedMargin;
  if (!c.tile_ppt && n) {
    uint32_t lg = 0;
    while ((n >> (lg + 1)) != 0) lg++;
    const uint32_t cls = lg / 2, ppt = cls < 16 ? srv->cfg.tile_ppt_by_class[cls] : 0u;
    if (ppt) c.tile_ppt = ppt;
  }
  if (n > c.capacity) return sfail(MPM_ERR_CAPACITY, "%u particles > capacity %u", n, c.capacity);
  int rc = validate_config(c);
  if (rc) return rc;
  auto r = std::make_unique<Request>();
  r->cfg = c;
  r->n = n;
  r->steps = steps;
  r->footprint = scene_footprint(c);
  if (r->footprint > srv->cfg.hbm_budget)
    return sfail(MPM_ERR_NOMEM, "scene needs %zu bytes, budget %llu", r->footprint,
                 (unsigned long long)srv->cfg.hbm_budget);
  r->pos3.assign(pos3, pos3 + (size_t)n * 3);
  if (vel3) r->vel3.assign(vel3, vel3 + (size_t)n * 3);
  if (C9) r->C9.assign(C9, C9 + (size_t)n * 9);
  if (mass) r->mass.assign(mass, mass + n);
  std::lock_guard<std::mutex> lk(srv->mu);
  if (srv->stop || srv->loop_error)
    return sfail(MPM_ERR_STATE, "server is stopped%s%s", srv->loop_msg.empty() ? "" : ": ",
                 srv->loop_msg.c_str());
  r->id = srv->next_id++;
  r->t_submit = srv->now();
  *id = r->id;
  srv->queue.push_back(r->id);
  srv->reqs[r->id] = std::move(r);
  srv->stats.submitted++;
  srv->cv_work.notify_all();
  return MPM_OK;
}

int mpm_server_poll(mpm_server *srv, uint64_t id, mpm_request_info *info) {
  if (!srv || !info) return sfail(MPM_ERR_INVALID, "mpm_server_poll: NULL argument");
  std::lock_guard<std::mutex> lk(srv->mu);
  auto it = srv->reqs.find(id);
  if (it == srv->reqs.end()) return sfail(MPM_ERR_NOT_FOUND, "request %llu unknown", (unsigned long long)id);
  const Request &r = *it->second;
  info->state = r.state;
  info->steps = r.steps;
  info->steps_done = r.done;
  info->n = r.n;
  info->t_submit = r.t_submit;
  info->t_admit = r.t_admit;
  info->t_done = r.t_done;
  info->footprint = r.footprint;
  info->tile_ppt = r.cfg.tile_ppt ? r.cfg.tile_ppt : 8u;
  info->pad = 0;
  return MPM_OK;
}

int mpm_server_wait(mpm_server *srv, uint64_t id, double timeout_s, int *state) {
  if (!srv) return sfail(MPM_ERR_INVALID, "mpm_server_wait: server is NULL");
  std::unique_lock<std::mutex> lk(srv->mu);
  auto deadline = Clock::now() + std::chrono::duration_cast<Clock::duration>(
                                     std::chrono::duration<double>(timeout_s < 0 ? 1e9 : timeout_s));
  for (;;) {
    auto it = srv->reqs.find(id);
    if (it == srv->reqs.end())
      return sfail(MPM_ERR_NOT_FOUND, "request %llu unknown", (unsigned long long)id);
    const int st = it->second->state;
    if (state) *state = st;
    if (st == MPM_REQ_DONE || st == MPM_REQ_FAILED || st == MPM_REQ_CANCELLED) return MPM_OK;
    if (srv->loop_error) return sfail(srv->loop_error, "server loop failed: %s", srv->loop_msg.c_str());
    if (srv->cv_done.wait_until(lk, deadline) == std::cv_status::timeout) {
      if (state) *state = srv->reqs.count(id) ? srv->reqs[id]->state : MPM_REQ_FAILED;
      return MPM_OK;
    }
  }
}

int mpm_server_result(mpm_server *srv, uint64_t id, float *pos4, float *vel4, float *C9,
                      float *density) {
  if (!srv) return sfail(MPM_ERR_INVALID, "mpm_server_result: server is NULL");
  std::unique_ptr<Request> r;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(id);
    if (it == srv->reqs.end())
      return sfail(MPM_ERR_NOT_FOUND, "request %llu unknown", (unsigned long long)id);
    const int st = it->second->state;
    if (st == MPM_REQ_FAILED) {
      std::string msg = it->second->error;
      srv->reqs.erase(it);
      return sfail(MPM_ERR_STATE, "request failed: %s", msg.c_str());
    }
    if (st == MPM_REQ_CANCELLED) {
      srv->reqs.erase(it);
      return sfail(MPM_ERR_STATE, "request was cancelled");
    }
    if (st != MPM_REQ_DONE) return sfail(MPM_ERR_STATE, "request not finished");
    // a finished request is the caller's alone: copied out after the lock
    r = std::move(it->second);
    srv->reqs.erase(it);
  }
  if (r->res) {
    const size_t n = r->n;
    if (pos4) memcpy(pos4, r->res, n * 16);
    if (vel4) memcpy(vel4, r->res + 4 * n, n * 16);
    if (C9) memcpy(C9, r->res + 8 * n, n * 36);
    if (density) memcpy(density, r->res + 17 * n, n * 4);
    srv->results->release(r->res);
  } else {
    if (pos4) memcpy(pos4, r->pos4.data(), r->pos4.size() * 4);
    if (vel4) memcpy(vel4, r->vel4.data(), r->vel4.size() * 4);
    if (C9) memcpy(C9, r->C9out.data(), r->C9out.size() * 4);
    if (density) memcpy(density, r->density.data(), r->density.size() * 4);
  }
  return MPM_OK;
}

int mpm_server_cancel(mpm_server *srv, uint64_t id) {
  if (!srv) return sfail(MPM_ERR_INVALID, "mpm_server_cancel: server is NULL");
  std::lock_guard<std::mutex> lk(srv->mu);
  auto it = srv->reqs.find(id);
  if (it == srv->reqs.end())
    return sfail(MPM_ERR_NOT_FOUND, "request %llu unknown", (unsigned long long)id);
  it->second->cancel = true;
  srv->cv_work.notify_all();
  return MPM_OK;
}

int mpm_server_get_stats(mpm_server *srv, mpm_server_stats *out) {
  if (!srv || !out) return sfail(MPM_ERR_INVALID, "mpm_server_get_stats: NULL argument");
  std::lock_guard<std::mutex> lk(srv->mu);
  *out = srv->stats;
  out->queued = (uint32_t)srv->queue.size();
  out->active = (uint32_t)srv->active.size();
  out->bytes_resident = srv->bytes_resident;
  out->elapsed_seconds = srv->now();
  return MPM_OK;
}

int mpm_server_step_records(mpm_server *srv, mpm_step_record *rec, uint32_t cap,
                            uint32_t *count) {
  if (!srv) return sfail(MPM_ERR_INVALID, "mpm_server_step_records: server is NULL");
  std::lock_guard<std::mutex> lk(srv->mu);
  const uint32_t n = (uint32_t)srv->step_rec.size();
  if (count) *count = n;
  if (rec)
    for (uint32_t i = 0; i < n && i < cap; i++) rec[i] = srv->step_rec[i];
  return MPM_OK;
}

int mpm_server_step_times(mpm_server *srv, float *ms, uint32_t cap, uint32_t *count) {
  if (!srv) return sfail(MPM_ERR_INVALID, "mpm_server_step_times: server is NULL");
  std::lock_guard<std::mutex> lk(srv->mu);
  const uint32_t n = (uint32_t)srv->step_ms.size();
  if (count) *count = n;
  if (ms)
    for (uint32_t i = 0; i < n && i < cap; i++) ms[i] = srv->step_ms[i];
  return MPM_OK;
}

}  // extern "C"
