// eqf_server.hip -- continuous batching of scene requests on one GPU.
//
// A server thread owns every admitted scene.  Each tick it launches ONE fused
// frame kernel (eqf_frames_cyclic, grid = resident wavefront-mode scenes) that
// advances every such scene by one frame (fluid.rs:437-524); red-black scenes
// (any size, for grids past the wavefront's 1026 limit) advance beside it by
// one frame of their launch group (every kernel covers all of them).  While the
// tick runs,
// the thread admits queued requests whose device memory fits the budget:
// it allocates them, writes Fluid::new's state, marks obstacles and uploads
// sources on each scene's own stream, so admission overlaps the running tick.
// Admitted scenes join at the next tick; finished scenes are read back and
// freed at tick boundaries.  Every scene runs its own workgroup with its own
// data, so a scene's result does not depend on which others share its ticks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "eqf_internal.h"

using namespace eqf;

namespace {

using Clock = std::chrono::steady_clock;

struct Req {
  uint64_t ticket = 0;
  eqf_params p{};
  int64_t nframes = 0;
  std::vector<int64_t> rects;
  std::vector<float> sd, su, sv, iu, iv;  // copied inputs
  // inputs read in place (borrow_inputs): the caller keeps them valid until admission
  const float *bsd = nullptr, *bsu = nullptr, *bsv = nullptr, *biu = nullptr, *biv = nullptr;
  bool keep_vel = false;
  int state = EQF_REQ_QUEUED;
  bool cancel = false;
  bool forget = false;  // released while queued or resident: dropped from the ticket map once retired
  int64_t frames_done = 0;
  double t_submit = 0, t_admit = 0, t_done = 0;
  eqf_scene* scene = nullptr;
  bool rb = false;  // red-black mode: own frame graph on the red-black stream
  uint64_t bytes = 0;
  int64_t cells = 0;
  std::vector<float> out_d, out_u, out_v;
};

}  // namespace

struct eqf_server {
  eqf_ctx* ctx = nullptr;
  eqf_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv;       // wakes the server thread
  std::condition_variable cv_idle;  // wakes eqf_server_wait_idle
  std::deque<Req*> queue;
  std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
  std::vector<Req*> active;
  uint64_t next_ticket = 1;
  uint64_t used_bytes = 0;
  uint64_t peak_bytes = 0;
  bool stop = false;
  std::thread th;
  hipStream_t stream = nullptr;   // ticks
  hipStream_t astream = nullptr;  // admissions (overlap the running tick)
  // red-black scenes' frames, beside the fused launch; two streams so that
  // the frames of two scenes overlap (small kernels, partly filled grids)
  static constexpr int kRbStreams = 2;
  hipStream_t rstream[kRbStreams] = {nullptr, nullptr};
  int rb_next = 0;
  // launch groups of the resident red-black scenes, one per (diffusion,
  // pressure) iteration counts: every frame kernel covers all scenes of a group
  std::map<std::pair<int, int>, void*> rb_groups;
  SceneDesc* d_descs = nullptr;
  int desc_cap = 0;
  Clock::time_point t0;
  // statistics
  int64_t ticks = 0, scene_frames = 0, max_resident = 0;
  double cell_updates = 0, busy_ms = 0;
  std::vector<std::pair<double, int>> tick_lat;  // (ms, resident scenes)
  std::vector<double> queue_lat;
  std::string thread_error;

  double now_ms() const {
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
  }
};

// Device bytes of a scene of this size and mode with frame sources.
static uint64_t estimate_bytes(const eqf_params& p) {
  const int S = (int)p.size, N = S - 2, ns = (N + 63) / 64, P = cyc_period(N);
  if (p.gs_mode == EQF_GS_REDBLACK) {
    // fields + sources, rule bytes, ping-pong field, row flags, break columns, lists
    const uint64_t fs = (uint64_t)rm_field_floats(S), cells = (uint64_t)S * S;
    return (NFIELDS + 3) * fs * 4 + cells + cells * 4 + (uint64_t)S * 5 + (uint64_t)N * 96 + sizeof(SceneDesc) +
           8192;
  }
  const uint64_t fs = (uint64_t)cyc_field_floats(S, P, ns);
  return (NFIELDS + 3) * fs * 4 + 3ull * (uint64_t)cyc_code_bytes(P, ns) + (uint64_t)S + (uint64_t)N * 96 +
         sizeof(SceneDesc) + 4096;
}

static int admit(eqf_server* sv, Req* r, std::string& err) {
  eqf_scene* s = nullptr;
  int rc = eqf_internal_scene_create(sv->ctx, &r->p, sv->astream, &s);
  for (size_t k = 0; !rc && k + 4 <= r->rects.size(); k += 4)
    rc = eqf_fill_rect(s, r->rects[k], r->rects[k + 1], r->rects[k + 2], r->rects[k + 3]);
  auto in = [](const std::vector<float>& v, const float* b) { return v.empty() ? b : v.data(); };
  const float *iu = in(r->iu, r->biu), *iv = in(r->iv, r->biv);
  const float *sd = in(r->sd, r->bsd), *su = in(r->su, r->bsu), *svv = in(r->sv, r->bsv);
  if (!rc && iu) rc = eqf_upload(s, EQF_VX, iu);
  if (!rc && iv) rc = eqf_upload(s, EQF_VY, iv);
  if (!rc && sd) rc = eqf_set_frame_source(s, EQF_DENSITY, sd);
  if (!rc && su) rc = eqf_set_frame_source(s, EQF_VX, su);
  if (!rc && svv) rc = eqf_set_frame_source(s, EQF_VY, svv);
  if (!rc) rc = eqf_internal_prepare(s);
  if (!rc) rc = eqf_sync(s);
  if (rc) {
    err = eqf_last_error();
    if (s) eqf_scene_destroy(s);
    return rc;
  }
  r->rb = s->layout == L_ROWMAJOR;
  if (r->rb) {  // its frames (and graph) live on a red-black stream
    s->stream = sv->rstream[sv->rb_next];
    sv->rb_next = (sv->rb_next + 1) % eqf_server::kRbStreams;
  }
  r->scene = s;
  r->cells = (int64_t)(r->p.size - 2) * (r->p.size - 2);
  // release the host copies of the inputs (borrowed ones are not read again)
  r->bsd = r->bsu = r->bsv = r->biu = r->biv = nullptr;
  std::vector<float>().swap(r->sd);
  std::vector<float>().swap(r->su);
  std::vector<float>().swap(r->sv);
  std::vector<float>().swap(r->iu);
  std::vector<float>().swap(r->iv);
  return EQF_OK;
}

static void finish(eqf_server* sv, Req* r) {
  const size_t n = (size_t)r->p.size * r->p.size;
  if (!r->cancel) {
    r->out_d.resize(n);
    if (eqf_download(r->scene, EQF_DENSITY, r->out_d.data()) != EQF_OK) r->state = EQF_REQ_FAILED;
    if (r->keep_vel) {
      r->out_u.resize(n);
      r->out_v.resize(n);
      (void)eqf_download(r->scene, EQF_VX, r->out_u.data());
      (void)eqf_download(r->scene, EQF_VY, r->out_v.data());
    }
  }
  eqf_scene_destroy(r->scene);
  r->scene = nullptr;
}

static void server_loop(eqf_server* sv) {
  (void)hipSetDevice(sv->ctx->device);
  std::vector<SceneDesc> descs;
  std::vector<uint64_t> last_uids;  // scenes (by uid: never reused) whose descriptors d_descs holds
  for (;;) {
    std::vector<Req*> act;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      sv->cv.wait(lk, [sv] { return sv->stop || !sv->queue.empty() || !sv->active.empty(); });
      if (sv->stop) break;
      // drop cancelled resident scenes
      for (auto it = sv->active.begin(); it != sv->active.end();) {
        Req* r = *it;
        if (r->cancel) {
          lk.unlock();
          finish(sv, r);
          lk.lock();
          sv->used_bytes -= r->bytes;
          if (r->state != EQF_REQ_FAILED) r->state = EQF_REQ_CANCELLED;
          it = sv->active.erase(it);
          if (r->forget) sv->reqs.erase(r->ticket);
        } else {
          ++it;
        }
      }
      act = sv->active;
    }
    // Launch one frame of every resident scene: the wavefront-mode scenes in
    // one fused launch (largest first), the red-black ones by their launch
    // groups (one per pair of iteration counts, one graph replay each).
    bool launched = false, launched_rb = false;
    double t_launch = sv->now_ms();
    std::vector<Req*> wf, rbs;
    for (Req* r : act) (r->rb ? rbs : wf).push_back(r);
    if (!wf.empty()) {
      std::stable_sort(wf.begin(), wf.end(), [](const Req* a, const Req* b) { return a->cells > b->cells; });
      int maxns = 1;
      for (Req* r : wf) maxns = std::max(maxns, r->scene->ns);
      std::vector<uint64_t> uids(wf.size());
      for (size_t i = 0; i < wf.size(); ++i) uids[i] = wf[i]->scene->uid;
      hipError_t e = hipSuccess;
      if (uids != last_uids) {
        descs.resize(wf.size());
        for (size_t i = 0; i < wf.size(); ++i) {
          descs[i] = wf[i]->scene->desc;
          descs[i].prio = eqf_batch_prio(wf[i]->scene->ns, maxns);
        }
        last_uids.clear();
        if ((int)wf.size() > sv->desc_cap) {
          const int cap = std::max<int>((int)wf.size(), 2 * sv->desc_cap);
          (void)hipFree(sv->d_descs);
          sv->d_descs = nullptr;
          sv->desc_cap = 0;
          e = hipMalloc(&sv->d_descs, sizeof(SceneDesc) * cap);
          if (e == hipSuccess) sv->desc_cap = cap;
          else sv->d_descs = nullptr;
        }
        if (e == hipSuccess)
          e = hipMemcpyAsync(sv->d_descs, descs.data(), sizeof(SceneDesc) * wf.size(), hipMemcpyHostToDevice,
                             sv->stream);
        if (e == hipSuccess) last_uids = uids;
      }
      int lds = 0, waves = 1;
      for (Req* r : wf) {
        lds = std::max(lds, cyc_lds_floats(r->scene->ns, r->scene->P));
        waves = std::max(waves, eqf_scene_waves(r->scene->ns, r->scene->P));
      }
      if (e == hipSuccess) e = eqf_internal_launch(sv->d_descs, (int)wf.size(), waves, lds, 1, sv->stream);
      if (e != hipSuccess) {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->thread_error = hipGetErrorString(e);
      } else {
        launched = true;
      }
    }
    {
      std::map<std::pair<int, int>, std::vector<Req*>> by_iters;
      for (Req* r : rbs) by_iters[{r->scene->desc.diff_iters, r->scene->desc.press_iters}].push_back(r);
      for (auto& [key, reqs] : by_iters) {
        std::stable_sort(reqs.begin(), reqs.end(), [](const Req* a, const Req* b) { return a->cells > b->cells; });
        std::vector<eqf_scene*> members;
        for (Req* r : reqs) members.push_back(r->scene);
        std::string err;
        void*& g = sv->rb_groups[key];
        if (!g) {
          g = rb_batch_create(members, sv->rstream[sv->rb_next], err);
          sv->rb_next = (sv->rb_next + 1) % eqf_server::kRbStreams;
        }
        int rc = g ? EQF_OK : EQF_EINVAL;
        if (g) {
          rb_batch_set_scenes(g, members);
          rc = rb_batch_step(g, 1, err);
        }
        if (rc != EQF_OK) {
          std::lock_guard<std::mutex> lk(sv->mu);
          sv->thread_error = err;
          for (Req* r : reqs) {
            r->state = EQF_REQ_FAILED;
            r->cancel = true;  // dropped at the next tick
          }
        } else {
          launched_rb = true;
        }
      }
    }
    auto tick_running = [&] {
      if (launched && hipStreamQuery(sv->stream) == hipErrorNotReady) return true;
      for (hipStream_t q : sv->rstream)
        if (launched_rb && hipStreamQuery(q) == hipErrorNotReady) return true;
      return false;
    };
    // Admit queued requests while the tick runs: at least one per tick, then
    // more only while the launch is still in flight (with nothing resident,
    // for up to 50 ms), so admission never delays the resident scenes' ticks.
    std::vector<Req*> admitted;
    const double t_adm0 = sv->now_ms();
    for (;;) {
      if (!admitted.empty()) {
        if ((launched || launched_rb) && !tick_running()) break;
        if (!launched && !launched_rb && sv->now_ms() - t_adm0 > 50.0) break;
      }
      Req* r = nullptr;
      {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (sv->queue.empty()) break;
        Req* h = sv->queue.front();
        if (h->cancel) {
          sv->queue.pop_front();
          h->state = EQF_REQ_CANCELLED;
          if (h->forget) sv->reqs.erase(h->ticket);
          continue;
        }
        const size_t resident = sv->active.size() + admitted.size();
        if (sv->cfg.max_active > 0 && resident >= (size_t)sv->cfg.max_active) break;
        if (sv->used_bytes + h->bytes > sv->cfg.hbm_budget_bytes) break;
        sv->queue.pop_front();
        sv->used_bytes += h->bytes;
        sv->peak_bytes = std::max(sv->peak_bytes, sv->used_bytes);
        r = h;
      }
      std::string err;
      if (admit(sv, r, err) != EQF_OK) {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->used_bytes -= r->bytes;
        r->state = EQF_REQ_FAILED;
        sv->thread_error = err;
        continue;
      }
      r->t_admit = sv->now_ms();
      admitted.push_back(r);
    }
    double tick_ms = 0;
    if (launched) (void)hipStreamSynchronize(sv->stream);
    if (launched_rb)
      for (hipStream_t q : sv->rstream) (void)hipStreamSynchronize(q);
    if (launched || launched_rb) tick_ms = sv->now_ms() - t_launch;
    // Retire finished scenes, then add the admitted ones.
    std::vector<Req*> done;
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      if (launched || launched_rb) {
        std::vector<Req*> ran = launched ? wf : std::vector<Req*>();
        for (Req* r : rbs)
          if (!r->cancel) ran.push_back(r);
        sv->ticks++;
        sv->scene_frames += (int64_t)ran.size();
        sv->busy_ms += tick_ms;
        sv->tick_lat.emplace_back(tick_ms, (int)ran.size());
        for (Req* r : ran) {
          r->frames_done++;
          sv->cell_updates += (double)r->cells;
        }
      }
      for (auto it = sv->active.begin(); it != sv->active.end();) {
        if ((*it)->frames_done >= (*it)->nframes) {
          done.push_back(*it);
          it = sv->active.erase(it);
        } else {
          ++it;
        }
      }
    }
    for (Req* r : done) finish(sv, r);
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      for (Req* r : done) {
        sv->used_bytes -= r->bytes;
        if (r->state != EQF_REQ_FAILED) r->state = r->cancel ? EQF_REQ_CANCELLED : EQF_REQ_DONE;
        r->t_done = sv->now_ms();
        if (r->forget) sv->reqs.erase(r->ticket);
      }
      for (Req* r : admitted) {
        // Staging buffers are only needed again for the final read-back; free
        // them now that no tick is running (hipFree waits for the device).
        eqf_dev_trim(r->scene);
        if (r->nframes <= 0) {  // nothing to run
          finish(sv, r);
          sv->used_bytes -= r->bytes;
          r->state = r->cancel ? EQF_REQ_CANCELLED : EQF_REQ_DONE;
          r->t_done = sv->now_ms();
          if (r->forget) sv->reqs.erase(r->ticket);
          continue;
        }
        r->state = EQF_REQ_ACTIVE;
        sv->active.push_back(r);
        sv->queue_lat.push_back(r->t_admit - r->t_submit);
      }
      sv->max_resident = std::max<int64_t>(sv->max_resident, (int64_t)sv->active.size());
      if (sv->queue.empty() && sv->active.empty()) sv->cv_idle.notify_all();
    }
  }
  // shutdown: free everything still resident
  std::lock_guard<std::mutex> lk(sv->mu);
  for (Req* r : sv->active) {
    eqf_scene_destroy(r->scene);
    r->scene = nullptr;
    r->state = EQF_REQ_CANCELLED;
  }
  sv->active.clear();
  sv->cv_idle.notify_all();
}

extern "C" int eqf_server_create(eqf_ctx* ctx, const eqf_server_config* cfg, eqf_server** out) {
  if (!ctx || !cfg || !out) return eqf_internal_error(EQF_EINVAL, "NULL argument");
  if (cfg->hbm_budget_bytes == 0) return eqf_internal_error(EQF_EINVAL, "hbm_budget_bytes must be > 0");
  if (cfg->max_active < 0) return eqf_internal_error(EQF_EINVAL, "max_active must be >= 0");
  if (hipSetDevice(ctx->device) != hipSuccess) return eqf_internal_error(EQF_EDEVICE, "hipSetDevice failed");
  auto* sv = new eqf_server;
  sv->ctx = ctx;
  sv->cfg = *cfg;
  sv->t0 = Clock::now();
  bool ok = hipStreamCreate(&sv->stream) == hipSuccess && hipStreamCreate(&sv->astream) == hipSuccess;
  for (hipStream_t& q : sv->rstream) ok = ok && hipStreamCreate(&q) == hipSuccess;
  if (!ok) {
    if (sv->stream) (void)hipStreamDestroy(sv->stream);
    if (sv->astream) (void)hipStreamDestroy(sv->astream);
    for (hipStream_t q : sv->rstream)
      if (q) (void)hipStreamDestroy(q);
    delete sv;
    return eqf_internal_error(EQF_EDEVICE, "hipStreamCreate failed");
  }
  sv->th = std::thread(server_loop, sv);
  *out = sv;
  return EQF_OK;
}

extern "C" int eqf_server_submit(eqf_server* sv, const eqf_request* req, uint64_t* ticket) {
  if (!sv || !req || !ticket) return eqf_internal_error(EQF_EINVAL, "NULL argument");
  if (req->nframes < 0) return eqf_internal_error(EQF_EINVAL, "nframes < 0");
  if (req->n_rects < 0 || (req->n_rects > 0 && !req->rects)) return eqf_internal_error(EQF_EINVAL, "bad rects");
  int rc = eqf_internal_check_params(&req->params);
  if (rc) return rc;
  auto r = std::make_unique<Req>();
  r->p = req->params;
  r->nframes = req->nframes;
  r->keep_vel = req->keep_velocity != 0;
  r->rects.assign(req->rects, req->rects + 4 * (size_t)req->n_rects);
  const size_t n = (size_t)req->params.size * req->params.size;
  auto cp = [n](const float* a, std::vector<float>& v) {
    if (a) v.assign(a, a + n);
  };
  if (req->borrow_inputs) {
    r->bsd = req->src_density;
    r->bsu = req->src_vx;
    r->bsv = req->src_vy;
    r->biu = req->init_vx;
    r->biv = req->init_vy;
  } else {
    cp(req->src_density, r->sd);
    cp(req->src_vx, r->su);
    cp(req->src_vy, r->sv);
    cp(req->init_vx, r->iu);
    cp(req->init_vy, r->iv);
  }
  r->bytes = estimate_bytes(r->p);
  std::lock_guard<std::mutex> lk(sv->mu);
  if (r->bytes > sv->cfg.hbm_budget_bytes)
    return eqf_internal_error(EQF_ENOMEM, "request does not fit the server's HBM budget");
  r->ticket = sv->next_ticket++;
  r->t_submit = sv->now_ms();
  *ticket = r->ticket;
  sv->queue.push_back(r.get());
  sv->reqs.emplace(r->ticket, std::move(r));
  sv->cv.notify_all();
  return EQF_OK;
}

extern "C" int eqf_server_poll(eqf_server* sv, uint64_t ticket, eqf_request_status* st) {
  if (!sv || !st) return eqf_internal_error(EQF_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->reqs.find(ticket);
  if (it == sv->reqs.end()) return eqf_internal_error(EQF_ENOTFOUND, "unknown ticket");
  const Req* r = it->second.get();
  st->state = r->state;
  st->reserved = 0;
  st->frames_done = r->frames_done;
  st->submit_ms = r->t_submit;
  st->admit_ms = r->t_admit;
  st->done_ms = r->t_done;
  return EQF_OK;
}

extern "C" int eqf_server_result(eqf_server* sv, uint64_t ticket, int field, float* out) {
  if (!sv || !out) return eqf_internal_error(EQF_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->reqs.find(ticket);
  if (it == sv->reqs.end()) return eqf_internal_error(EQF_ENOTFOUND, "unknown ticket");
  const Req* r = it->second.get();
  if (r->state != EQF_REQ_DONE) return eqf_internal_error(EQF_EBUSY, "request has not finished");
  const std::vector<float>* v = field == EQF_DENSITY ? &r->out_d : field == EQF_VX ? &r->out_u : field == EQF_VY ? &r->out_v : nullptr;
  if (!v || v->empty()) return eqf_internal_error(EQF_EINVAL, "field not kept for this request");
  std::memcpy(out, v->data(), v->size() * sizeof(float));
  return EQF_OK;
}

extern "C" int eqf_server_release(eqf_server* sv, uint64_t ticket) {
  if (!sv) return eqf_internal_error(EQF_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->reqs.find(ticket);
  if (it == sv->reqs.end()) return eqf_internal_error(EQF_ENOTFOUND, "unknown ticket");
  Req* r = it->second.get();
  if (r->state == EQF_REQ_QUEUED || r->state == EQF_REQ_ACTIVE) {
    r->cancel = true;  // the server thread retires it and forgets the ticket
    r->forget = true;
    sv->cv.notify_all();
    return EQF_OK;
  }
  sv->reqs.erase(it);
  return EQF_OK;
}

extern "C" int eqf_server_wait_idle(eqf_server* sv, double timeout_ms) {
  if (!sv) return eqf_internal_error(EQF_EINVAL, "NULL argument");
  std::unique_lock<std::mutex> lk(sv->mu);
  auto idle = [sv] { return sv->queue.empty() && sv->active.empty(); };
  if (timeout_ms < 0) {
    sv->cv_idle.wait(lk, idle);
  } else if (!sv->cv_idle.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), idle)) {
    return eqf_internal_error(EQF_EBUSY, "server still busy at timeout");
  }
  if (!sv->thread_error.empty()) return eqf_internal_error(EQF_EDEVICE, sv->thread_error);
  return EQF_OK;
}

static double pct_weighted(std::vector<std::pair<double, int>> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  int64_t total = 0;
  for (auto& x : v) total += x.second;
  const double target = q * (double)(total - 1);
  int64_t acc = 0;
  for (auto& x : v) {
    acc += x.second;
    if ((double)acc > target) return x.first;
  }
  return v.back().first;
}

extern "C" int eqf_server_stats_get(eqf_server* sv, eqf_server_stats* st) {
  if (!sv || !st) return eqf_internal_error(EQF_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  std::memset(st, 0, sizeof(*st));
  st->ticks = sv->ticks;
  st->scene_frames = sv->scene_frames;
  st->cell_updates = sv->cell_updates;
  st->busy_ms = sv->busy_ms;
  st->p50_frame_ms = pct_weighted(sv->tick_lat, 0.50);
  st->p99_frame_ms = pct_weighted(sv->tick_lat, 0.99);
  std::vector<std::pair<double, int>> q;
  for (double x : sv->queue_lat) q.emplace_back(x, 1);
  st->p50_queue_ms = pct_weighted(q, 0.50);
  st->p99_queue_ms = pct_weighted(q, 0.99);
  st->max_resident = sv->max_resident;
  st->peak_bytes = sv->peak_bytes;
  return EQF_OK;
}

extern "C" int eqf_server_destroy(eqf_server* sv) {
  if (!sv) return EQF_OK;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
    sv->cv.notify_all();
  }
  if (sv->th.joinable()) sv->th.join();
  (void)hipStreamSynchronize(sv->stream);
  for (auto& kv : sv->rb_groups) {
    for (hipStream_t q : sv->rstream) (void)hipStreamSynchronize(q);
    rb_batch_destroy(kv.second);
  }
  for (hipStream_t q : sv->rstream) (void)hipStreamSynchronize(q);
  (void)hipFree(sv->d_descs);
  (void)hipStreamDestroy(sv->stream);
  (void)hipStreamDestroy(sv->astream);
  for (hipStream_t q : sv->rstream) (void)hipStreamDestroy(q);
  delete sv;
  return EQF_OK;
}
