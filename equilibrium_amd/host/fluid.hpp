// fluid.hpp -- C++ mirror of the reference simulation API (src/simulation of
// vkabadzhova/equilibrium) over the C ABI of libeqfluid.so.  Names and
// argument meanings follow the Rust items cited on each declaration; where
// the reference panics, this header throws std::invalid_argument /
// std::runtime_error with the reference's message.  Header-only; link with
// -leqfluid.
#pragma once
#include <algorithm>
#include <array>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/eqfluid.h"

namespace equilibrium {

// fluid.rs:12-17
enum class ContainerWall : uint8_t { DefaultWall = EQF_DEFAULT_WALL, NoWall = EQF_NO_WALL };

// configs.rs:5-33
struct SimulationConfigs {
  float delta_t = 0.02f;
  int64_t frames = 16;
  uint32_t size = 128;
  static SimulationConfigs new_(float delta_t, int64_t frames, uint32_t size) { return {delta_t, frames, size}; }
};

using Rgba = std::array<uint8_t, 4>;  // egui::Color32 as its four bytes

// configs.rs:37-60
struct FluidConfigs {
  float diffusion = 0.0f;
  float viscousity = 0.001f;
  bool has_perlin_noise = true;
  Rgba fluid_color{208, 88, 157, 220};
  Rgba world_color{94, 146, 162, 128};
};

using Point = std::pair<int64_t, int64_t>;  // line_drawing::Point<i64>

// obstacle.rs:29-98
class Rectangle {
 public:
  Point down_left_point, up_right_point;
  Rectangle(Point dl, Point ur, uint32_t fluid_container_size) : down_left_point(dl), up_right_point(ur) {
    approximate_points_ = {dl, ur};
    if (!are_all_points_valid(fluid_container_size)) throw std::invalid_argument("Invalid input for Rectangle");
  }
  static Rectangle default_() { return Rectangle({80, 80}, {110, 110}, SimulationConfigs{}.size); }
  bool are_all_points_valid(int64_t size) const {
    return eqf_rect_validate(down_left_point.first, down_left_point.second, up_right_point.first,
                             up_right_point.second, (uint32_t)size) == EQF_OK;
  }
  std::vector<Point>& get_approximate_points() { return approximate_points_; }

 private:
  std::vector<Point> approximate_points_;
};

// obstacle.rs:12-27 (one variant)
struct ObstaclesType {
  Rectangle rectangle;
  std::vector<Point>& get_approximate_points() { return rectangle.get_approximate_points(); }
};

namespace detail {
inline void check(int rc) {
  if (rc != EQF_OK) throw std::runtime_error(std::string("eqf: ") + eqf_last_error());
}
inline eqf_ctx* default_ctx() {
  static eqf_ctx* ctx = [] {
    eqf_ctx* c = nullptr;
    check(eqf_ctx_create(0, &c));
    return c;
  }();
  return ctx;
}
}  // namespace detail

// fluid.rs:51-620.  The six fields live on the GPU; the pub fields of the
// reference are read through density() / velocities_x() / velocities_y() /
// cells_type(), which download the current state.
class Fluid {
 public:
  FluidConfigs fluid_configs;
  SimulationConfigs simulation_configs;

  // fluid.rs:93 (gs_mode: EQF_GS_WAVEFRONT reproduces the reference bit for bit)
  Fluid(FluidConfigs f, SimulationConfigs s, int gs_mode = EQF_GS_WAVEFRONT, int init_times = 1)
      : fluid_configs(f), simulation_configs(s) {
    eqf_params p;
    eqf_params_default(&p);
    p.size = s.size;
    p.delta_t = s.delta_t;
    p.frames = s.frames;
    p.diffusion = f.diffusion;
    p.viscosity = f.viscousity;
    p.gs_mode = gs_mode;
    p.init_times = init_times;
    eqf_scene* h = nullptr;
    detail::check(eqf_scene_create(detail::default_ctx(), &p, &h));
    scene_.reset(h);
  }
  static Fluid new_(FluidConfigs f, SimulationConfigs s) { return Fluid(f, s); }
  // fluid.rs:83-89: Default runs init a second time
  static Fluid default_() { return Fluid(FluidConfigs{}, SimulationConfigs{}, EQF_GS_WAVEFRONT, 2); }
  // #[derive(Clone)] (fluid.rs:51)
  Fluid clone() const {
    eqf_scene* h = nullptr;
    detail::check(eqf_scene_clone(scene_.get(), &h));
    return Fluid(fluid_configs, simulation_configs, h);
  }

  // fluid.rs:112-117
  static std::pair<int64_t, int64_t> to_coordinate(int64_t idx, int64_t size) { return {idx % size, idx / size}; }

  // fluid.rs:120-131
  void add_density(uint32_t x, uint32_t y, float amount) { detail::check(eqf_add_density(scene_.get(), x, y, amount)); }
  void add_velocity(uint32_t x, uint32_t y, float ax, float ay) {
    detail::check(eqf_add_velocity(scene_.get(), x, y, ax, ay));
  }

  // fluid.rs:575-599 with the Perlin angle and the thread_rng point supplied
  // by the caller (the noise/rand/geo crates are not reproduced)
  void add_noise(float angle_degrees, uint32_t rand_x, uint32_t rand_y) {
    detail::check(eqf_add_noise(scene_.get(), angle_degrees, rand_x, rand_y));
  }

  // fluid.rs:437-524
  void step() {
    sync_params();
    detail::check(eqf_step(scene_.get(), 1));
  }

  // fluid.rs:610-619
  void fill_obstacle(ObstaclesType& obstacle) {
    auto& pts = obstacle.get_approximate_points();
    detail::check(eqf_fill_rect(scene_.get(), pts[0].first, pts[0].second, pts[1].first, pts[1].second));
  }

  std::vector<float> density() const { return field(EQF_DENSITY); }
  std::vector<float> velocities_x() const { return field(EQF_VX); }
  std::vector<float> velocities_y() const { return field(EQF_VY); }
  std::vector<ContainerWall> cells_type() const {
    std::vector<uint8_t> m((size_t)simulation_configs.size * simulation_configs.size);
    detail::check(eqf_get_mask(scene_.get(), m.data()));
    std::vector<ContainerWall> out(m.size());
    for (size_t i = 0; i < m.size(); ++i) out[i] = (ContainerWall)m[i];
    return out;
  }
  std::vector<float> field(int f) const {
    std::vector<float> v((size_t)simulation_configs.size * simulation_configs.size);
    detail::check(eqf_download(scene_.get(), f, v.data()));
    return v;
  }
  eqf_scene* handle() const { return scene_.get(); }

 private:
  struct Del {
    void operator()(eqf_scene* s) const { eqf_scene_destroy(s); }
  };
  std::unique_ptr<eqf_scene, Del> scene_;
  Fluid(FluidConfigs f, SimulationConfigs s, eqf_scene* h) : fluid_configs(f), simulation_configs(s), scene_(h) {}
  void sync_params() {
    eqf_params p;
    detail::check(eqf_scene_get_params(scene_.get(), &p));
    if (p.delta_t != simulation_configs.delta_t || p.frames != simulation_configs.frames ||
        p.diffusion != fluid_configs.diffusion || p.viscosity != fluid_configs.viscousity) {
      p.delta_t = simulation_configs.delta_t;
      p.frames = simulation_configs.frames;
      p.diffusion = fluid_configs.diffusion;
      p.viscosity = fluid_configs.viscousity;
      detail::check(eqf_scene_set_params(scene_.get(), &p));
    }
  }
};

// renderer_helpers.rs:87-165.  render_image returns the RGBA8 pixels
// (x + y*size, four bytes each) computed on the GPU instead of a PNG file.
struct RenderingListener {
  Rgba obstacles_color{255, 0, 0, 255};  // egui::Color32::RED

  eqf_colors colors(const Fluid& f) const {
    eqf_colors c;
    for (int k = 0; k < 4; ++k) {
      c.fluid[k] = f.fluid_configs.fluid_color[k];
      c.world[k] = f.fluid_configs.world_color[k];
      c.obstacle[k] = obstacles_color[k];
    }
    return c;
  }
  std::vector<uint8_t> render_image(const Fluid& f) const {
    const eqf_colors c = colors(f);
    std::vector<uint8_t> px((size_t)4 * f.simulation_configs.size * f.simulation_configs.size);
    detail::check(eqf_render_rgba(f.handle(), &c, px.data()));
    return px;
  }
};

// One frame handed to the consumer: density (row-major) and, when asked
// for, the rendered RGBA8 image.  Both point into a pinned ring slot that
// stays valid until the FrameSnapshots object moves on.
struct FrameView {
  int64_t frame;
  const float* density;
  const uint8_t* rgba;
};

// Ring of per-frame host copies of one Fluid (eqf_snapshot_*).
class FrameSnapshots {
 public:
  FrameSnapshots(Fluid& f, int depth, const eqf_colors* colors) {
    eqf_snapshot* h = nullptr;
    detail::check(eqf_snapshot_create(f.handle(), depth, colors != nullptr, colors, &h));
    h_.reset(h);
  }
  void capture(int64_t frame) { detail::check(eqf_snapshot_capture(h_.get(), frame)); }
  FrameView acquire() {
    FrameView v{};
    detail::check(eqf_snapshot_acquire(h_.get(), &v.frame, &v.density, &v.rgba));
    return v;
  }
  void release() { detail::check(eqf_snapshot_release(h_.get())); }

 private:
  struct Del {
    void operator()(eqf_snapshot* s) const { eqf_snapshot_destroy(s); }
  };
  std::unique_ptr<eqf_snapshot, Del> h_;
};

// renderer_helpers.rs:29-81 without the mpsc channel: on_frame(i, fluid)
// receives each frame's state.
struct CurrentSimulation {
  Fluid fluid = Fluid::default_();
  std::vector<ObstaclesType> obstacles{ObstaclesType{Rectangle::default_()}};

  void mark_fluid_obstacles() {
    for (auto& o : obstacles) fluid.fill_obstacle(o);
  }
  template <class F>
  void simulate(F&& on_frame) {
    mark_fluid_obstacles();
    for (int64_t i = 0; i < fluid.simulation_configs.frames; ++i) {
      fluid.step();
      on_frame(i, fluid);
    }
  }
  // The same frames, each handed over as a FrameView whose copy to the host
  // ran behind the following steps (up to `depth` frames in flight);
  // listener != nullptr adds its rendered image.
  template <class F>
  void simulate_snapshots(F&& on_frame, int depth = 2, const RenderingListener* listener = nullptr) {
    mark_fluid_obstacles();
    eqf_colors c{};
    if (listener) c = listener->colors(fluid);
    FrameSnapshots snaps(fluid, depth, listener ? &c : nullptr);
    const int64_t frames = fluid.simulation_configs.frames;
    auto hand_over = [&] {
      on_frame(snaps.acquire());
      snaps.release();
    };
    for (int64_t i = 0; i < frames; ++i) {
      if (i >= depth) hand_over();
      fluid.step();
      snaps.capture(i);
    }
    for (int64_t i = 0; i < std::min<int64_t>(depth, frames); ++i) hand_over();
  }
};

}  // namespace equilibrium
