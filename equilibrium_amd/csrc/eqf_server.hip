// eqf_server.hip -- continuous batching of scene requests on one GPU.
//
// A server thread owns every admitted scene.  Resident scenes are split into
// lanes that tick independently, each on its own stream: small wavefront-mode
// scenes (at most 4 strips of 64 rows, N <= 256), large wavefront-mode scenes
// (5-16 strips), and red-black scenes (any size).  A lane's tick advances its
// scenes by one frame (fluid.rs:437-524): one fused launch of
// eqf_frames_cyclic over the lane's scenes, sized by the lane's own largest
// scene, or one frame graph per red-black launch group (one group per pair of
// iteration counts).  When a lane's tick ends, the thread retires the lane's
// finished scenes, adds the ones admitted to it meanwhile and launches its
// next tick, so a small scene's frame takes its own lane's time rather than
// the largest resident scene's.  Within a wavefront lane's tick, a scene
// whose frame chain is k times shorter than the lane's longest runs up to k
// frames (at most 16, never past its request) in its workgroup, so the
// lane's scenes advance at their own pace.  Between those events the thread admits
// queued requests while their device memory fits the budget (allocation,
// Fluid::new's state, obstacles, sources) on a fourth stream, overlapping the
// running ticks.  Every scene runs its own workgroups with its own data, so a
// scene's result does not depend on which others share its ticks.
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
  int tick_frames = 1;  // frames of the tick in flight
  double t_submit = 0, t_admit = 0, t_done = 0;
  eqf_scene* scene = nullptr;
  bool rb = false;  // red-black mode: advanced by its launch group on the red-black lane
  int lane = 0;
  uint64_t bytes = 0;
  int64_t cells = 0;
  std::vector<float> out_d, out_u, out_v;
};

}  // namespace

namespace {
// Lanes of resident scenes (see the file comment).  With the admission stream
// that is four streams, the number of hardware queues HIP gives a process by
// default: streams sharing a queue would run their ticks one after the other.
enum { LANE_SMALL = 0, LANE_LARGE = 1, LANE_RB = 2, NLANES = 3 };
constexpr int kSmallMaxStrips = 4;

struct Lane {
  hipStream_t st = nullptr;
  hipEvent_t ev = nullptr;          // recorded after each tick
  std::vector<Req*> members;        // resident scenes of the lane
  std::vector<Req*> ticking;        // scenes of the tick in flight
  bool busy = false;
  double t_launch = 0;
  // wavefront lanes: the descriptors of the scenes last launched
  SceneDesc* d_descs = nullptr;
  int desc_cap = 0;
  std::vector<uint64_t> last_uids;  // (scene uid (never reused), frames per tick) pairs that d_descs holds
  std::vector<SceneDesc> descs;
};

int lane_of(const eqf_scene* s) {
  if (s->layout == L_ROWMAJOR) return LANE_RB;
  return s->ns <= kSmallMaxStrips ? LANE_SMALL : LANE_LARGE;
}
}  // namespace

struct eqf_server {
  eqf_ctx* ctx = nullptr;
  eqf_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv_tick;  // wakes the tick thread
  std::condition_variable cv_work;  // wakes the admission thread
  std::condition_variable cv_idle;  // wakes eqf_server_wait_idle
  std::deque<Req*> queue;
  std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
  std::vector<Req*> active;         // admitted and not yet retired
  std::vector<Req*> incoming[NLANES];  // admitted, joining their lane at its next tick
  std::vector<Req*> retiring;       // left their lane, to be read back and freed
  uint64_t next_ticket = 1;
  uint64_t used_bytes = 0;
  uint64_t peak_bytes = 0;
  bool stop = false;
  std::thread th_tick, th_work;
  hipStream_t astream = nullptr;  // admissions (overlap the running ticks)
  // the resident scenes' device memory, allocated and freed in astream's order
  hipMemPool_t pool = nullptr;
  Lane lanes[NLANES];
  // launch groups of the resident red-black scenes, one per (diffusion,
  // pressure) iteration counts, on the red-black lane's stream: every frame
  // kernel covers all scenes of a group
  std::map<std::pair<int, int>, void*> rb_groups;
  Clock::time_point t0;
  // statistics
  int64_t ticks = 0, scene_frames = 0, max_resident = 0;
  double cell_updates = 0, busy_ms = 0;
  int busy_lanes = 0;
  double busy_since = 0;
  std::vector<std::pair<uint32_t, float>> frame_lat;   // (size, tick ms) per scene-frame
  std::vector<std::pair<uint32_t, double>> queue_lat;  // (size, submit -> admit ms) per request
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
    return (NFIELDS + 3) * fs * 4 + (uint64_t)rm_code_bytes(S) + cells * 4 + (uint64_t)S * 5 + (uint64_t)N * 96 +
           sizeof(SceneDesc) + 8192;
  }
  const uint64_t fs = (uint64_t)cyc_field_floats(S, P, ns);
  return (NFIELDS + 3) * fs * 4 + 3ull * (uint64_t)cyc_code_bytes(P, ns) + (uint64_t)S + (uint64_t)N * 96 +
         sizeof(SceneDesc) + 4096;
}

static int admit(eqf_server* sv, Req* r, std::string& err) {
  eqf_scene* s = nullptr;
  int rc = eqf_internal_scene_create(sv->ctx, &r->p, sv->astream, sv->pool, &s);
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
  r->lane = lane_of(s);
  // the scene keeps the admission stream for its own operations (its
  // read-back at retirement, after its lane's last tick has ended); its
  // frames run on its lane's stream
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

// A lane's tick has ended (called with sv->mu held).
static void lane_tick_done(eqf_server* sv, Lane& L, double now) {
  const float ms = (float)(now - L.t_launch);
  sv->ticks++;
  for (Req* r : L.ticking) {
    const int f = r->tick_frames;
    r->frames_done += f;
    sv->scene_frames += f;
    sv->cell_updates += (double)r->cells * f;
    for (int k = 0; k < f; ++k) sv->frame_lat.emplace_back(r->p.size, ms / f);
  }
  L.ticking.clear();
  L.busy = false;
  if (--sv->busy_lanes == 0) sv->busy_ms += now - sv->busy_since;
}

// Dependent wavefront steps of one frame of a scene: its lin_solves' chains
// (sweeps over m streams per strip, strip and stream fill); the one-pass
// phases add a few passes over the strip.
static double frame_chain(const eqf_scene* s, int sched) {
  const int W = eqf_scene_waves(s->ns, s->P, sched);
  const int lag = gs_stream_lag(sched);
  const int m = cyc_streams(W, s->ns, s->P, sched);
  auto solve = [&](int it) {
    if (it <= 0) return 0.0;
    const int mm = std::min(m, it);
    return (double)((it + mm - 1) / mm) * s->P + (double)(s->ns - 1) * gs_strip_lag(sched) +
           (double)(mm - 1) * lag + 64.0;
  };
  return 3.0 * solve(s->desc.diff_iters) + 2.0 * solve(s->desc.press_iters) + 4.0 * s->P;
}

// Launches one tick of every scene of an idle lane.  Returns an error
// message, or "" on success.
static std::string lane_launch(eqf_server* sv, int li) {
  Lane& L = sv->lanes[li];
  L.ticking = L.members;
  hipError_t e = hipSuccess;
  std::string err;
  if (li != LANE_RB) {
    auto& wf = L.ticking;
    std::stable_sort(wf.begin(), wf.end(), [](const Req* a, const Req* b) { return a->cells > b->cells; });
    int maxns = 1, lds = 0, waves = 1;
    double chain_max = 0;
    int bigP = 16;  // the period of the largest scene (most strips)
    for (Req* r : wf) {
      if (r->scene->ns > maxns || (r->scene->ns == maxns && r->scene->P > bigP)) bigP = r->scene->P;
      maxns = std::max(maxns, r->scene->ns);
    }
    const int sched = eqf_pick_sched(maxns, bigP, 16);
    for (Req* r : wf) {
      waves = std::max(waves, eqf_scene_waves(r->scene->ns, r->scene->P, sched));
      chain_max = std::max(chain_max, frame_chain(r->scene, sched));
    }
    for (Req* r : wf) lds = std::max(lds, cyc_lds_floats(r->scene->ns, r->scene->P, waves, sched));
    std::vector<uint64_t> uids(2 * wf.size());
    for (size_t i = 0; i < wf.size(); ++i) {
      Req* r = wf[i];
      const int64_t k = (int64_t)(chain_max / frame_chain(r->scene, sched));
      r->tick_frames = (int)std::max<int64_t>(1, std::min<int64_t>({k, 16, r->nframes - r->frames_done}));
      uids[2 * i] = r->scene->uid;
      uids[2 * i + 1] = (uint64_t)r->tick_frames;
    }
    if (uids != L.last_uids) {
      L.descs.resize(wf.size());
      int maxP = 0;
      for (Req* r : wf) maxP = std::max(maxP, r->scene->P);
      for (size_t i = 0; i < wf.size(); ++i) {
        L.descs[i] = wf[i]->scene->desc;
        L.descs[i].prio = eqf_batch_prio(wf[i]->scene->ns, maxns);
        L.descs[i].frame_mult = wf[i]->tick_frames;
        L.descs[i].stream_l2 = eqf_stream_l2(maxP, (int)wf.size());
      }
      L.last_uids.clear();
      if ((int)wf.size() > L.desc_cap) {
        // the lane is idle: nothing reads the old array any more
        const int cap = std::max<int>((int)wf.size(), 2 * L.desc_cap);
        (void)hipFree(L.d_descs);
        L.d_descs = nullptr;
        L.desc_cap = 0;
        e = hipMalloc(&L.d_descs, sizeof(SceneDesc) * cap);
        if (e == hipSuccess) L.desc_cap = cap;
      }
      // pageable source: the copy has read it when the call returns
      if (e == hipSuccess)
        e = hipMemcpyAsync(L.d_descs, L.descs.data(), sizeof(SceneDesc) * wf.size(), hipMemcpyHostToDevice, L.st);
      if (e == hipSuccess) L.last_uids = uids;
    }
    if (e == hipSuccess) e = eqf_internal_launch(L.d_descs, (int)wf.size(), waves, lds, 1, sched, L.st);
    if (e != hipSuccess) err = hipGetErrorString(e);
  } else {
    std::map<std::pair<int, int>, std::vector<eqf_scene*>> by_iters;
    std::stable_sort(L.ticking.begin(), L.ticking.end(),
                     [](const Req* a, const Req* b) { return a->cells > b->cells; });
    for (Req* r : L.ticking) {
      r->tick_frames = 1;
      by_iters[{r->scene->desc.diff_iters, r->scene->desc.press_iters}].push_back(r->scene);
    }
    for (auto& [key, members] : by_iters) {
      void*& g = sv->rb_groups[key];
      if (!g) g = rb_batch_create(members, L.st, err);
      if (!g) break;
      rb_batch_set_scenes(g, members);
      if (rb_batch_step(g, 1, err) != EQF_OK) break;
    }
  }
  if (err.empty()) {
    e = hipEventRecord(L.ev, L.st);
    if (e != hipSuccess) err = hipGetErrorString(e);
  }
  if (!err.empty()) {
    L.ticking.clear();
    return err;
  }
  std::lock_guard<std::mutex> lk(sv->mu);
  L.busy = true;
  L.t_launch = sv->now_ms();
  if (sv->busy_lanes++ == 0) sv->busy_since = L.t_launch;
  return "";
}

// Admits the queue's head if its memory fits; true if one was admitted (or
// dropped).
static bool admit_one(eqf_server* sv) {
  Req* r = nullptr;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    while (!sv->queue.empty() && sv->queue.front()->cancel) {
      Req* h = sv->queue.front();
      sv->queue.pop_front();
      h->state = EQF_REQ_CANCELLED;
      if (h->forget) sv->reqs.erase(h->ticket);
    }
    if (sv->queue.empty()) return false;
    Req* h = sv->queue.front();
    if (sv->cfg.max_active > 0 && sv->active.size() >= (size_t)sv->cfg.max_active) return false;
    if (sv->used_bytes + h->bytes > sv->cfg.hbm_budget_bytes) return false;
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
    return true;
  }
  r->t_admit = sv->now_ms();
  if (r->nframes <= 0) {  // nothing to run
    finish(sv, r);
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->used_bytes -= r->bytes;
    r->state = r->cancel ? EQF_REQ_CANCELLED : EQF_REQ_DONE;
    r->t_done = sv->now_ms();
    if (r->forget) sv->reqs.erase(r->ticket);
    return true;
  }
  std::lock_guard<std::mutex> lk(sv->mu);
  r->state = EQF_REQ_ACTIVE;
  sv->active.push_back(r);
  sv->incoming[r->lane].push_back(r);
  sv->queue_lat.emplace_back(r->p.size, r->t_admit - r->t_submit);
  sv->max_resident = std::max<int64_t>(sv->max_resident, (int64_t)sv->active.size());
  sv->cv_tick.notify_all();
  return true;
}

// Tick thread: detects ended ticks, moves an idle lane's finished or
// cancelled scenes to the retiring list, adds the scenes admitted to it and
// launches its next tick.  It does no host copies, so a lane restarts within
// about 100 us of its tick's end whatever the admissions are doing.
static void tick_loop(eqf_server* sv) {
  (void)hipSetDevice(sv->ctx->device);
  for (;;) {
    bool progress = false;
    for (Lane& L : sv->lanes) {
      if (!L.busy) continue;
      const hipError_t q = hipEventQuery(L.ev);
      if (q == hipErrorNotReady) continue;
      std::lock_guard<std::mutex> lk(sv->mu);
      if (q != hipSuccess) sv->thread_error = hipGetErrorString(q);
      lane_tick_done(sv, L, sv->now_ms());
      progress = true;
    }
    for (int li = 0; li < NLANES; ++li) {
      Lane& L = sv->lanes[li];
      if (L.busy) continue;
      {
        std::lock_guard<std::mutex> lk(sv->mu);
        bool moved = false;
        for (auto it = L.members.begin(); it != L.members.end();) {
          Req* r = *it;
          if (r->cancel || r->frames_done >= r->nframes) {
            sv->retiring.push_back(r);
            it = L.members.erase(it);
            moved = true;
          } else {
            ++it;
          }
        }
        if (moved) sv->cv_work.notify_all();
        L.members.insert(L.members.end(), sv->incoming[li].begin(), sv->incoming[li].end());
        sv->incoming[li].clear();
      }
      if (L.members.empty()) continue;
      const std::string err = lane_launch(sv, li);
      if (!err.empty()) {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->thread_error = err;
        for (Req* r : L.members) {
          r->state = EQF_REQ_FAILED;
          r->cancel = true;  // retired at the next pass
        }
      }
      progress = true;
    }
    std::unique_lock<std::mutex> lk(sv->mu);
    if (sv->stop) break;
    if (!progress) {
      bool any = false;
      for (const Lane& L : sv->lanes) any = any || L.busy || !L.members.empty();
      for (const auto& v : sv->incoming) any = any || !v.empty();
      if (any)  // poll the running ticks
        sv->cv_tick.wait_for(lk, std::chrono::microseconds(100));
      else
        sv->cv_tick.wait(lk, [sv] {
          if (sv->stop) return true;
          for (const auto& v : sv->incoming)
            if (!v.empty()) return true;
          return false;
        });
    }
  }
}

// Admission thread: reads back and frees retired scenes, then admits queued
// requests one at a time while their memory fits the budget.
static void work_loop(eqf_server* sv) {
  (void)hipSetDevice(sv->ctx->device);
  for (;;) {
    std::vector<Req*> ret;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      auto ready = [sv] {
        if (sv->stop || !sv->retiring.empty()) return true;
        if (sv->queue.empty()) return false;
        const Req* h = sv->queue.front();
        if (h->cancel) return true;
        if (sv->cfg.max_active > 0 && sv->active.size() >= (size_t)sv->cfg.max_active) return false;
        return sv->used_bytes + h->bytes <= sv->cfg.hbm_budget_bytes;
      };
      if (sv->queue.empty() && sv->active.empty()) sv->cv_idle.notify_all();
      sv->cv_work.wait(lk, ready);
      if (sv->stop) break;
      ret.swap(sv->retiring);
    }
    for (Req* r : ret) finish(sv, r);
    if (!ret.empty()) {
      std::lock_guard<std::mutex> lk(sv->mu);
      for (Req* r : ret) {
        // leaves `active` only with its final state, so eqf_server_wait_idle
        // never returns while a request still reads as resident
        sv->active.erase(std::find(sv->active.begin(), sv->active.end(), r));
        sv->used_bytes -= r->bytes;
        if (r->state != EQF_REQ_FAILED) r->state = r->cancel ? EQF_REQ_CANCELLED : EQF_REQ_DONE;
        r->t_done = sv->now_ms();
        if (r->forget) sv->reqs.erase(r->ticket);
      }
    }
    admit_one(sv);
  }
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
  bool ok = hipStreamCreate(&sv->astream) == hipSuccess;
  if (ok) {
    hipMemPoolProps props;
    std::memset(&props, 0, sizeof(props));
    props.allocType = hipMemAllocationTypePinned;
    props.handleTypes = hipMemHandleTypeNone;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = ctx->device;
    ok = hipMemPoolCreate(&sv->pool, &props) == hipSuccess;
    // keep freed memory in the pool for the next admissions (the budget bounds it)
    uint64_t keep = UINT64_MAX;
    ok = ok && hipMemPoolSetAttribute(sv->pool, hipMemPoolAttrReleaseThreshold, &keep) == hipSuccess;
  }
  for (Lane& L : sv->lanes) {
    ok = ok && hipStreamCreate(&L.st) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&L.ev, hipEventDisableTiming) == hipSuccess;
  }
  if (!ok) {
    if (sv->pool) (void)hipMemPoolDestroy(sv->pool);
    if (sv->astream) (void)hipStreamDestroy(sv->astream);
    for (Lane& L : sv->lanes) {
      if (L.st) (void)hipStreamDestroy(L.st);
      if (L.ev) (void)hipEventDestroy(L.ev);
    }
    delete sv;
    return eqf_internal_error(EQF_EDEVICE, "hipStreamCreate failed");
  }
  sv->th_tick = std::thread(tick_loop, sv);
  sv->th_work = std::thread(work_loop, sv);
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
  sv->cv_work.notify_all();
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
    r->cancel = true;  // the server threads retire it and forget the ticket
    r->forget = true;
    sv->cv_work.notify_all();
    sv->cv_tick.notify_all();
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

// Statistics over the scenes of sizes [lo, hi] (every scene: 0, UINT32_MAX).
static void stats_for(eqf_server* sv, uint32_t lo, uint32_t hi, eqf_server_stats* st) {
  std::memset(st, 0, sizeof(*st));
  std::vector<std::pair<double, int>> f, q;
  double cells = 0;
  for (auto& x : sv->frame_lat)
    if (x.first >= lo && x.first <= hi) {
      f.emplace_back(x.second, 1);
      cells += (double)(x.first - 2) * (double)(x.first - 2);
    }
  for (auto& x : sv->queue_lat)
    if (x.first >= lo && x.first <= hi) q.emplace_back(x.second, 1);
  st->ticks = sv->ticks;
  st->scene_frames = (int64_t)f.size();
  st->cell_updates = cells;
  // wall time with at least one lane running, the current stretch included
  st->busy_ms = sv->busy_ms + (sv->busy_lanes > 0 ? sv->now_ms() - sv->busy_since : 0.0);
  st->p50_frame_ms = pct_weighted(f, 0.50);
  st->p99_frame_ms = pct_weighted(f, 0.99);
  st->p50_queue_ms = pct_weighted(q, 0.50);
  st->p99_queue_ms = pct_weighted(q, 0.99);
  st->max_resident = sv->max_resident;
  st->peak_bytes = sv->peak_bytes;
}

extern "C" int eqf_server_stats_get(eqf_server* sv, eqf_server_stats* st) {
  if (!sv || !st) return eqf_internal_error(EQF_EINVAL, "NULL argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  stats_for(sv, 0, UINT32_MAX, st);
  return EQF_OK;
}

extern "C" int eqf_server_stats_sizes(eqf_server* sv, uint32_t min_size, uint32_t max_size, eqf_server_stats* st) {
  if (!sv || !st) return eqf_internal_error(EQF_EINVAL, "NULL argument");
  if (min_size > max_size) return eqf_internal_error(EQF_EINVAL, "min_size > max_size");
  std::lock_guard<std::mutex> lk(sv->mu);
  stats_for(sv, min_size, max_size, st);
  return EQF_OK;
}

extern "C" int eqf_server_destroy(eqf_server* sv) {
  if (!sv) return EQF_OK;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
    sv->cv_tick.notify_all();
    sv->cv_work.notify_all();
  }
  if (sv->th_tick.joinable()) sv->th_tick.join();
  if (sv->th_work.joinable()) sv->th_work.join();
  // wait for the ticks in flight, then free every scene still admitted
  for (Lane& L : sv->lanes)
    if (L.busy) (void)hipEventSynchronize(L.ev);
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    for (Req* r : sv->active) {
      eqf_scene_destroy(r->scene);
      r->scene = nullptr;
      r->state = EQF_REQ_CANCELLED;
    }
    sv->active.clear();
    sv->retiring.clear();
    for (auto& v : sv->incoming) v.clear();
    sv->cv_idle.notify_all();
  }
  for (Lane& L : sv->lanes) (void)hipStreamSynchronize(L.st);
  for (auto& kv : sv->rb_groups) rb_batch_destroy(kv.second);
  for (Lane& L : sv->lanes) {
    (void)hipFree(L.d_descs);
    (void)hipStreamDestroy(L.st);
    (void)hipEventDestroy(L.ev);
  }
  (void)hipStreamSynchronize(sv->astream);
  (void)hipStreamDestroy(sv->astream);
  (void)hipMemPoolDestroy(sv->pool);
  delete sv;
  return EQF_OK;
}
