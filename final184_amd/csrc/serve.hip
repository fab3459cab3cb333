// serve.hip — the serving loop: continuous batching of voxelize / mips / trace requests
// over scenes kept resident in device memory under a byte budget.
//
// Threads: submit() (any caller thread) appends to the waiting queue; one scheduler
// thread forms batches; one worker thread per slot issues its batch on the slot's own HIP
// stream and waits for it. Every decision is taken under the server mutex by the
// scheduler (admission, eviction, pinning) or by a worker finishing its batch (unpin,
// completions); device work and the blocking waits happen outside it.
//
// Batch formation (each time a slot is idle and something changed): walk the waiting
// requests in arrival order and take each one that can run now:
//   - its scene is resident with the same grid and neither rebuild is asked for nor a
//     build in flight: trace only (pins the scene);
//   - otherwise it needs a build: not while another in-flight batch uses the scene, and
//     only once per scene per batch; the residency cache admits it (evicting unpinned
//     scenes least-recently-used, recycling a victim of the same grid shape) or says wait;
//     a scene larger than the whole budget fails the request.
// A batch is one voxelize batch + one mips batch over its builds, then one trace batch over
// every frame of every request in it, in stream order, so a request's trace sees its own
// build. A trace-only request never shares a scene with an in-flight build, so results do
// not depend on scheduling.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "residency.hpp"
#include "vgi_internal.hpp"

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Req {
  vgi_request r{};
  std::vector<vgi_mesh> meshes;
  std::vector<vgi_frame> frames;
  double start = 0.0;
  bool build = false;
};

struct Batch {
  std::vector<Req*> reqs;
  std::vector<uint64_t> pinned;  // one key per request
  std::vector<uint64_t> built;   // keys built by this batch
  std::vector<vgi_voxelize_item> vox;
  std::vector<vgi_mips_item> mips;
  std::vector<vgi_trace_item> trace;
};

struct Slot {
  void* stream = nullptr;
  std::thread th;
  std::condition_variable cv;
  bool has_work = false;
  bool busy = false;
  Batch batch;
};

bool same_grid(const vgi_grid& a, const vgi_grid& b) {
  return a.n == b.n && a.levels == b.levels && a.extent == b.extent && a.origin[0] == b.origin[0] &&
         a.origin[1] == b.origin[1] && a.origin[2] == b.origin[2];
}

}  // namespace

struct vgi_server {
  vgi_context* ctx = nullptr;
  vgi_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv_sched;  // submissions, completions, stop
  std::condition_variable cv_done;   // completions
  std::deque<Req*> waiting;
  std::deque<vgi_completion> done;
  vgi::Residency<vgi_scene*> res{0};
  std::unordered_map<uint64_t, vgi_grid> grid_of;
  std::vector<std::unique_ptr<Slot>> slots;
  std::thread sched;
  bool stop = false;
  uint64_t epoch = 0;  // bumps on every submission and completion
  uint32_t inflight = 0;
  uint64_t outstanding = 0;
  vgi_server_stats stats{};
  std::vector<vgi_scene*> to_release;
};

namespace {

void complete(vgi_server* S, Req* q, int status, double t_done) {
  vgi_completion c{};
  c.tag = q->r.tag;
  c.status = status;
  c.voxelized = q->build && status == VGI_OK ? 1u : 0u;
  c.arrival_s = q->r.arrival_s;
  c.start_s = q->start;
  c.done_s = t_done;
  S->done.push_back(c);
  S->stats.completed++;
  if (status != VGI_OK) S->stats.failed++;
  S->outstanding--;
  delete q;
}

// Fills one batch from the waiting queue (server mutex held).
void form_batch(vgi_server* S, Batch& b) {
  std::unordered_set<uint64_t> in_batch, built_here;
  uint64_t pixels = 0;
  for (auto it = S->waiting.begin(); it != S->waiting.end();) {
    if (b.reqs.size() >= S->cfg.max_batch_requests) break;
    if (S->cfg.max_batch_pixels && !b.reqs.empty() && pixels >= S->cfg.max_batch_pixels) break;
    Req* q = *it;
    const uint64_t key = q->r.scene_key;
    auto* e = S->res.find(key);
    if (e && e->building && !built_here.count(key)) {  // its build is in flight elsewhere
      ++it;
      continue;
    }
    const bool need_build = !e || q->r.revoxelize || !same_grid(S->grid_of[key], q->r.grid);
    if (need_build) {
      // one build per scene per batch, and never under another in-flight batch's reads
      if (in_batch.count(key) || (e && e->pins > 0)) {
        ++it;
        continue;
      }
      std::unordered_set<uint64_t> protect = in_batch;
      const vgi::SceneShape shape{q->r.grid.n, q->r.grid.levels};
      const uint64_t bytes = vgi_scene_bytes_estimate(&q->r.grid, S->cfg.precision, 0);
      auto ar = S->res.admit(key, shape, bytes, protect);
      if (ar.status == vgi::Residency<vgi_scene*>::Admit::kOverBudget) {
        it = S->waiting.erase(it);
        complete(S, q, VGI_E_BUDGET, now_s());
        continue;
      }
      if (ar.status == vgi::Residency<vgi_scene*>::Admit::kWait) {
        ++it;
        continue;
      }
      for (vgi_scene* h : ar.release) S->to_release.push_back(h);
      for (uint64_t k : ar.evicted) S->grid_of.erase(k);
      S->stats.evictions += ar.evicted.size();
      vgi_scene* sc = ar.handle;
      if (!sc) {
        if (vgi_scene_create(S->ctx, &sc) != VGI_OK) {
          S->res.remove(key, nullptr);
          it = S->waiting.erase(it);
          complete(S, q, VGI_E_NOMEM, now_s());
          continue;
        }
        S->res.set_handle(key, sc);
      }
      e = S->res.find(key);
      e->building = true;
      S->grid_of[key] = q->r.grid;
      q->build = true;
      built_here.insert(key);
      b.built.push_back(key);
      b.vox.push_back(vgi_voxelize_item{sc, q->meshes.data(), (uint32_t)q->meshes.size(), q->r.grid});
      b.mips.push_back(vgi_mips_item{sc, &q->r.lights});
    }
    e = S->res.find(key);
    e->pins++;
    S->res.touch(key);
    in_batch.insert(key);
    b.pinned.push_back(key);
    for (const vgi_frame& f : q->frames) {
      b.trace.push_back(vgi_trace_item{e->handle, f.gbuffer, f.cones, f.out_diffuse, f.out_specular});
      pixels += (uint64_t)f.gbuffer.width * f.gbuffer.height;
      S->stats.frames++;
      S->stats.pixels += (uint64_t)f.gbuffer.width * f.gbuffer.height;
    }
    b.reqs.push_back(q);
    it = S->waiting.erase(it);
  }
}

void worker(vgi_server* S, Slot* slot) {
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    slot->cv.wait(lk, [&] { return slot->has_work || S->stop; });
    if (!slot->has_work) return;  // stop with nothing assigned
    Batch b = std::move(slot->batch);
    slot->batch = Batch{};
    slot->has_work = false;
    lk.unlock();
    const double t0 = now_s();
    for (Req* q : b.reqs) q->start = t0;
    int rc = VGI_OK;
    if (!b.vox.empty()) rc = vgi_voxelize_batch(S->ctx, b.vox.data(), (uint32_t)b.vox.size(), slot->stream);
    if (rc == VGI_OK && !b.mips.empty())
      rc = vgi_build_mips_batch(S->ctx, b.mips.data(), (uint32_t)b.mips.size(), S->cfg.precision, slot->stream);
    if (rc == VGI_OK && !b.trace.empty())
      rc = vgi_trace_batch(S->ctx, b.trace.data(), (uint32_t)b.trace.size(), slot->stream);
    const int rs = vgi_sync(S->ctx, slot->stream);
    if (rc == VGI_OK) rc = rs;
    const double t1 = now_s();
    lk.lock();
    for (uint64_t k : b.pinned)
      if (auto* e = S->res.find(k)) e->pins--;
    for (uint64_t k : b.built) {
      auto* e = S->res.find(k);
      if (!e) continue;
      e->building = false;
      if (rc != VGI_OK) {  // a failed build leaves no usable scene
        vgi_scene* h = nullptr;
        if (e->pins == 0 && S->res.remove(k, &h)) {
          S->grid_of.erase(k);
          if (h) S->to_release.push_back(h);
        }
      }
    }
    S->stats.builds += rc == VGI_OK ? b.built.size() : 0;
    for (Req* q : b.reqs) complete(S, q, rc, t1);
    slot->busy = false;
    S->inflight--;
    S->epoch++;
    S->cv_sched.notify_all();
    S->cv_done.notify_all();
  }
}

void scheduler(vgi_server* S) {
  std::unique_lock<std::mutex> lk(S->mu);
  uint64_t seen = ~0ull;
  for (;;) {
    S->cv_sched.wait(lk, [&] {
      if (S->stop) return true;
      if (S->epoch == seen || S->waiting.empty()) return false;
      for (auto& s : S->slots)
        if (!s->busy) return true;
      return false;
    });
    if (S->stop) return;
    seen = S->epoch;
    for (auto& s : S->slots) {
      if (s->busy) continue;
      Batch b;
      form_batch(S, b);
      if (b.reqs.empty()) break;
      S->stats.batches++;
      S->stats.max_batch_requests_seen = std::max<uint32_t>(S->stats.max_batch_requests_seen, (uint32_t)b.reqs.size());
      s->batch = std::move(b);
      s->busy = true;
      s->has_work = true;
      S->inflight++;
      S->stats.max_concurrent_batches = std::max(S->stats.max_concurrent_batches, S->inflight);
      s->cv.notify_one();
    }
    if (!S->to_release.empty()) {  // evicted scenes of another shape (rare): free outside the lock
      std::vector<vgi_scene*> rel;
      rel.swap(S->to_release);
      lk.unlock();
      for (vgi_scene* h : rel) vgi_scene_release(h);
      lk.lock();
    }
    S->cv_done.notify_all();  // failed admissions may have completed requests
  }
}

}  // namespace

extern "C" {

double vgi_now(void) { return now_s(); }

int vgi_server_create(vgi_context* ctx, const vgi_server_config* cfg, vgi_server** out) {
  if (!ctx || !cfg || !out) return vgi::set_error(VGI_E_INVALID, "NULL argument");
  *out = nullptr;
  if (cfg->slots < 1 || cfg->slots > 16) return vgi::set_error(VGI_E_INVALID, "slots %u not in [1,16]", cfg->slots);
  if (cfg->max_batch_requests < 1) return vgi::set_error(VGI_E_INVALID, "max_batch_requests must be >= 1");
  if (cfg->precision > VGI_PREC_FP16X2) return vgi::set_error(VGI_E_INVALID, "precision %u", cfg->precision);
  if (cfg->memory_budget == 0) return vgi::set_error(VGI_E_INVALID, "memory_budget is 0");
  auto S = std::make_unique<vgi_server>();
  S->ctx = ctx;
  S->cfg = *cfg;
  S->res = vgi::Residency<vgi_scene*>(cfg->memory_budget);
  for (uint32_t i = 0; i < cfg->slots; i++) {
    auto s = std::make_unique<Slot>();
    int rc = vgi_stream_create(ctx, &s->stream);
    if (rc != VGI_OK) {
      for (auto& p : S->slots) vgi_stream_destroy(ctx, p->stream);
      return rc;
    }
    S->slots.push_back(std::move(s));
  }
  vgi_server* raw = S.release();
  for (auto& s : raw->slots) s->th = std::thread(worker, raw, s.get());
  raw->sched = std::thread(scheduler, raw);
  *out = raw;
  return VGI_OK;
}

int vgi_server_submit(vgi_server* S, const vgi_request* req) {
  if (!S || !req) return vgi::set_error(VGI_E_INVALID, "NULL argument");
  int rc = vgi::check_request_input(req->meshes, req->num_meshes, &req->grid, &req->lights, req->frames,
                                    req->num_frames);
  if (rc != VGI_OK) return rc;
  Req* q = new Req();
  q->r = *req;
  q->meshes.assign(req->meshes, req->meshes + req->num_meshes);
  q->frames.assign(req->frames, req->frames + req->num_frames);
  q->r.meshes = nullptr;
  q->r.frames = nullptr;
  if (q->r.arrival_s <= 0.0) q->r.arrival_s = now_s();
  std::lock_guard<std::mutex> lk(S->mu);
  if (S->stop) {
    delete q;
    return vgi::set_error(VGI_E_STATE, "server is stopping");
  }
  S->waiting.push_back(q);
  S->outstanding++;
  S->stats.submitted++;
  S->epoch++;
  S->cv_sched.notify_all();
  return VGI_OK;
}

int vgi_server_poll(vgi_server* S, vgi_completion* out, uint32_t max, uint32_t* count, uint32_t wait_us) {
  if (!S || !count || (max && !out)) return vgi::set_error(VGI_E_INVALID, "NULL argument");
  std::unique_lock<std::mutex> lk(S->mu);
  if (S->done.empty() && wait_us)
    S->cv_done.wait_for(lk, std::chrono::microseconds(wait_us), [&] { return !S->done.empty(); });
  uint32_t n = 0;
  while (n < max && !S->done.empty()) {
    out[n++] = S->done.front();
    S->done.pop_front();
  }
  *count = n;
  return VGI_OK;
}

int vgi_server_drain(vgi_server* S) {
  if (!S) return vgi::set_error(VGI_E_INVALID, "NULL argument");
  std::unique_lock<std::mutex> lk(S->mu);
  S->cv_done.wait(lk, [&] { return S->outstanding == 0; });
  return VGI_OK;
}

int vgi_server_get_stats(vgi_server* S, vgi_server_stats* out) {
  if (!S || !out) return vgi::set_error(VGI_E_INVALID, "NULL argument");
  std::lock_guard<std::mutex> lk(S->mu);
  *out = S->stats;
  out->resident_scenes = S->res.size();
  out->resident_bytes = S->res.used();
  out->peak_resident_bytes = S->res.peak();
  return VGI_OK;
}

int vgi_server_destroy(vgi_server* S) {
  if (!S) return VGI_OK;
  vgi_server_drain(S);
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
    S->cv_sched.notify_all();
    for (auto& s : S->slots) s->cv.notify_all();
  }
  S->sched.join();
  for (auto& s : S->slots) s->th.join();
  std::vector<vgi_scene*> rel;
  rel.swap(S->to_release);
  S->res.for_each([&](uint64_t, vgi::Residency<vgi_scene*>::Entry& e) {
    if (e.handle) rel.push_back(e.handle);
  });
  for (vgi_scene* h : rel) vgi_scene_release(h);
  for (auto& s : S->slots) vgi_stream_destroy(S->ctx, s->stream);
  delete S;
  return VGI_OK;
}

}  // extern "C"
