// The reference's own unit tests, restated against the C++ mirror of its API
// (equilibrium_amd/host/fluid.hpp) over libeqfluid.so, plus a step-parity
// check against the CPU oracle, and the batch and serving entry points
// against solo runs.  `--cpu` runs the tests that need no GPU.
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "../../equilibrium_amd/host/fluid.hpp"

using namespace equilibrium;

extern "C" {
// oracle/eqf_oracle.c (test infrastructure, loaded only by this test)
void ora_init_f32(float*, float*, float*, float*, float*, float*, uint8_t*, uint32_t, int);
void ora_fill_rect(uint8_t*, uint32_t, int64_t, int64_t, int64_t, int64_t);
void ora_run_f32(float*, float*, float*, float*, float*, float*, int64_t, float, float, float, int64_t, int64_t,
                 const uint8_t*, int, const float*, const float*, const float*, int64_t);
void ora_render_rgba(const float*, const uint8_t*, int64_t, const uint8_t*, const uint8_t*, const uint8_t*, uint8_t*);
}

static int failures = 0;
#define CHECK(cond)                                                  \
  do {                                                               \
    if (!(cond)) {                                                   \
      std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond); \
      ++failures;                                                    \
    }                                                                \
  } while (0)

// fluid.rs:626-635
static void to_coordinate() {
  const std::pair<int64_t, int64_t> coordinate{3, 4};
  const int64_t size = 10;
  const int64_t idx = coordinate.first + coordinate.second * size;  // idx!(3, 4, 10)
  CHECK(Fluid::to_coordinate(idx, size) == coordinate);
}

// obstacle.rs:100-115
static void rectangle_panics() {
  bool thrown = false;
  try {
    Rectangle({50, 120}, {127, 110}, 128);
  } catch (const std::invalid_argument& e) {
    thrown = std::strcmp(e.what(), "Invalid input for Rectangle") == 0;
  }
  CHECK(thrown);
  thrown = false;
  try {
    Rectangle({12, 12}, {10, 10}, 128);
  } catch (const std::invalid_argument&) {
    thrown = true;
  }
  CHECK(thrown);
  Rectangle r = Rectangle::default_();
  CHECK(r.get_approximate_points().size() == 2);
}

// renderer_helpers.rs:222-252
static void mark_fluid_obstacles_default_current_simulation() {
  CurrentSimulation sim;
  sim.mark_fluid_obstacles();
  int64_t walls = 0;
  for (auto c : sim.fluid.cells_type()) walls += c == ContainerWall::DefaultWall;
  const int64_t size = sim.fluid.simulation_configs.size;
  const int64_t image_parameter = 2 * (size + (size - 2));
  CHECK(walls == 30 * 30 + image_parameter);  // 1408
}

// Default simulation (size 128, dt 0.02, 16 frames = 16 sweeps per lin_solve,
// default rectangle) run for 4 frames on the GPU == the CPU oracle, bit for bit.
static void default_simulation_matches_oracle() {
  CurrentSimulation sim;
  sim.fluid.simulation_configs.frames = 4;
  std::vector<std::vector<float>> got;
  sim.simulate([&](int64_t, Fluid& f) { got = {f.density(), f.velocities_x(), f.velocities_y()}; });
  const uint32_t S = 128;
  std::vector<float> D(S * S), Sc(S * S), U(S * S), V(S * S), U0(S * S), V0(S * S);
  std::vector<uint8_t> m(S * S);
  ora_init_f32(D.data(), Sc.data(), U.data(), V.data(), U0.data(), V0.data(), m.data(), S, 2);
  ora_fill_rect(m.data(), S, 80, 80, 110, 110);
  ora_run_f32(D.data(), Sc.data(), U.data(), V.data(), U0.data(), V0.data(), S, 0.02f, 0.0f, 0.001f, 4, 4, m.data(), 0,
              nullptr, nullptr, nullptr, 4);
  CHECK(std::memcmp(got[0].data(), D.data(), D.size() * 4) == 0);
  CHECK(std::memcmp(got[1].data(), U.data(), U.size() * 4) == 0);
  CHECK(std::memcmp(got[2].data(), V.data(), V.size() * 4) == 0);
}

// The same simulation through simulate_snapshots with a RenderingListener:
// every frame's snapshot (density and RGBA image) == the oracle's state after
// that frame and its rendering, bit for bit, in frame order.
static void snapshots_match_oracle() {
  CurrentSimulation sim;
  sim.fluid.simulation_configs.frames = 4;
  RenderingListener listener;
  const uint32_t S = 128;
  std::vector<int64_t> frames;
  std::vector<std::vector<float>> dens;
  std::vector<std::vector<uint8_t>> imgs;
  sim.simulate_snapshots(
      [&](const FrameView& v) {
        frames.push_back(v.frame);
        dens.emplace_back(v.density, v.density + S * S);
        imgs.emplace_back(v.rgba, v.rgba + 4 * S * S);
      },
      2, &listener);
  CHECK(frames.size() == 4);
  std::vector<float> D(S * S), Sc(S * S), U(S * S), V(S * S), U0(S * S), V0(S * S);
  std::vector<uint8_t> m(S * S), img(4 * S * S);
  ora_init_f32(D.data(), Sc.data(), U.data(), V.data(), U0.data(), V0.data(), m.data(), S, 2);
  ora_fill_rect(m.data(), S, 80, 80, 110, 110);
  const FluidConfigs fc;
  for (size_t i = 0; i < frames.size(); ++i) {
    CHECK(frames[i] == (int64_t)i);
    ora_run_f32(D.data(), Sc.data(), U.data(), V.data(), U0.data(), V0.data(), S, 0.02f, 0.0f, 0.001f, 4, 4,
                m.data(), 0, nullptr, nullptr, nullptr, 1);
    ora_render_rgba(D.data(), m.data(), S, fc.fluid_color.data(), fc.world_color.data(),
                    listener.obstacles_color.data(), img.data());
    CHECK(std::memcmp(dens[i].data(), D.data(), D.size() * 4) == 0);
    CHECK(std::memcmp(imgs[i].data(), img.data(), img.size()) == 0);
  }
}

// The batch and serving entry points from C++ through the raw C ABI: three
// scenes of different sizes (1, 2 and 5 strips) with per-frame density and
// velocity sources and an obstacle, advanced (a) alone, (b) in one batch,
// (c) as server requests; all three give the same fields bit for bit.
static std::vector<float> ramp(uint32_t S, float scale) {
  std::vector<float> v((size_t)S * S);
  for (size_t i = 0; i < v.size(); ++i) v[i] = scale * (float)((i * 7919u) % 97u) / 97.0f;
  return v;
}

static eqf_scene* make_scene(uint32_t S, const std::vector<float>& sd, const std::vector<float>& su) {
  eqf_params p;
  eqf_params_default(&p);
  p.size = S;
  p.delta_t = 0.1f;
  p.viscosity = 1e-4f;
  p.diffusion = 1e-5f;
  p.gs_iters = 6;
  eqf_scene* h = nullptr;
  detail::check(eqf_scene_create(detail::default_ctx(), &p, &h));
  detail::check(eqf_fill_rect(h, S / 4, S / 4, S / 4 + 5, S / 4 + 7));
  detail::check(eqf_set_frame_source(h, EQF_DENSITY, sd.data()));
  detail::check(eqf_set_frame_source(h, EQF_VX, su.data()));
  return h;
}

static void batch_and_server_match_solo() {
  const uint32_t sizes[3] = {50, 100, 300};
  const int64_t frames = 3;
  std::vector<std::vector<float>> solo_d, batch_d, server_d;
  std::vector<std::vector<float>> sd, su;
  for (uint32_t S : sizes) {
    sd.push_back(ramp(S, 2.0f));
    su.push_back(ramp(S, 0.5f));
  }
  for (int i = 0; i < 3; ++i) {  // (a) alone
    eqf_scene* h = make_scene(sizes[i], sd[i], su[i]);
    detail::check(eqf_step(h, frames));
    std::vector<float> d((size_t)sizes[i] * sizes[i]);
    detail::check(eqf_download(h, EQF_DENSITY, d.data()));
    solo_d.push_back(d);
    detail::check(eqf_scene_destroy(h));
  }
  // (b) one batch, under each wavefront schedule (N=298 runs 2 / 3 / 3 sweep
  // streams per strip under the wide / ring / fine schedule, N=98 1 / 2 / 3)
  std::vector<std::vector<float>> ring_d, fine_d;
  for (int sched : {EQF_SCHED_WIDE, EQF_SCHED_RING, EQF_SCHED_FINE}) {
    eqf_scene* hs[3];
    for (int i = 0; i < 3; ++i) hs[i] = make_scene(sizes[i], sd[i], su[i]);
    eqf_batch* b = nullptr;
    detail::check(eqf_batch_create(detail::default_ctx(), hs, 3, &b));
    detail::check(eqf_batch_set_schedule(b, sched));
    detail::check(eqf_batch_step(b, frames));
    detail::check(eqf_batch_sync(b));
    double ms = 0;
    detail::check(eqf_batch_last_ms(b, &ms));
    CHECK(ms > 0);
    detail::check(eqf_batch_destroy(b));
    for (int i = 0; i < 3; ++i) {
      std::vector<float> d((size_t)sizes[i] * sizes[i]);
      detail::check(eqf_download(hs[i], EQF_DENSITY, d.data()));
      (sched == EQF_SCHED_RING ? ring_d : sched == EQF_SCHED_FINE ? fine_d : batch_d).push_back(d);
      detail::check(eqf_scene_destroy(hs[i]));
    }
  }
  {  // (c) server requests
    eqf_server_config cfg{};
    cfg.hbm_budget_bytes = 4ull << 30;
    eqf_server* sv = nullptr;
    detail::check(eqf_server_create(detail::default_ctx(), &cfg, &sv));
    uint64_t tickets[3];
    for (int i = 0; i < 3; ++i) {
      eqf_request rq{};
      eqf_params_default(&rq.params);
      rq.params.size = sizes[i];
      rq.params.delta_t = 0.1f;
      rq.params.viscosity = 1e-4f;
      rq.params.diffusion = 1e-5f;
      rq.params.gs_iters = 6;
      rq.nframes = frames;
      const int64_t rect[4] = {sizes[i] / 4, sizes[i] / 4, sizes[i] / 4 + 5, sizes[i] / 4 + 7};
      rq.n_rects = 1;
      rq.rects = rect;
      rq.src_density = sd[i].data();
      rq.src_vx = su[i].data();
      detail::check(eqf_server_submit(sv, &rq, &tickets[i]));  // inputs copied at submit
    }
    detail::check(eqf_server_wait_idle(sv, 60000.0));
    for (int i = 0; i < 3; ++i) {
      eqf_request_status st{};
      detail::check(eqf_server_poll(sv, tickets[i], &st));
      CHECK(st.state == EQF_REQ_DONE && st.frames_done == frames);
      std::vector<float> d((size_t)sizes[i] * sizes[i]);
      detail::check(eqf_server_result(sv, tickets[i], EQF_DENSITY, d.data()));
      server_d.push_back(d);
      detail::check(eqf_server_release(sv, tickets[i]));
    }
    eqf_server_stats st{};
    detail::check(eqf_server_stats_get(sv, &st));
    CHECK(st.scene_frames == 3 * frames);
    detail::check(eqf_server_destroy(sv));
  }
  for (int i = 0; i < 3; ++i) {
    CHECK(std::memcmp(solo_d[i].data(), batch_d[i].data(), solo_d[i].size() * 4) == 0);
    CHECK(std::memcmp(solo_d[i].data(), ring_d[i].data(), solo_d[i].size() * 4) == 0);
    CHECK(std::memcmp(solo_d[i].data(), fine_d[i].data(), solo_d[i].size() * 4) == 0);
    CHECK(std::memcmp(solo_d[i].data(), server_d[i].data(), solo_d[i].size() * 4) == 0);
  }
}

int main(int argc, char** argv) {
  const bool gpu = !(argc > 1 && std::strcmp(argv[1], "--cpu") == 0);
  to_coordinate();
  rectangle_panics();
  if (gpu) {
    mark_fluid_obstacles_default_current_simulation();
    default_simulation_matches_oracle();
    snapshots_match_oracle();
    batch_and_server_match_solo();
  }
  std::printf("%s: %d failure(s)\n", gpu ? "cpu+gpu" : "cpu", failures);
  return failures ? 1 : 0;
}
