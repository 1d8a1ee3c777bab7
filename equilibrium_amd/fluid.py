"""Python mirror of the reference simulation API (src/simulation), backed by
libeqfluid.so on the GPU.

Names and argument meanings follow the Rust items they mirror:

* ``SimulationConfigs`` -- configs.rs:5-33
* ``FluidConfigs``      -- configs.rs:37-60 (colours kept as plain tuples)
* ``ContainerWall``     -- fluid.rs:12-17
* ``Rectangle``         -- obstacle.rs:29-98 (invalid input raises
  ``ValueError("Invalid input for Rectangle")`` where the reference panics)
* ``ObstaclesType``     -- obstacle.rs:12-27
* ``Fluid``             -- fluid.rs:51-620; pub fields become properties that
  read the device-resident state, ``step`` runs on the GPU.
* ``CurrentSimulation`` -- renderer_helpers.rs:29-81 frame loop (without the
  mpsc channel: ``simulate`` calls a callback per frame; ``simulate_snapshots``
  yields per-frame host snapshots copied behind the following steps).
* ``RenderingListener`` -- renderer_helpers.rs:87-165: ``render_image`` builds
  the RGBA image on the device and returns it as a (size, size, 4) array
  instead of writing a PNG file.
"""
import enum
import math

import numpy as np

from . import _abi


class ContainerWall(enum.IntEnum):
    DefaultWall = _abi.DEFAULT_WALL
    NoWall = _abi.NO_WALL


class SimulationConfigs:
    def __init__(self, delta_t=0.02, frames=16, size=128):
        self.delta_t = float(delta_t)
        self.frames = int(frames)
        self.size = int(size)

    @classmethod
    def new(cls, delta_t, frames, fluid_container_size):
        return cls(delta_t, frames, fluid_container_size)

    @classmethod
    def default(cls):
        return cls()


class FluidConfigs:
    def __init__(self, diffusion=0.0, viscousity=0.001, has_perlin_noise=True,
                 fluid_color=(208, 88, 157, 220), world_color=(94, 146, 162, 128)):
        self.diffusion = float(diffusion)
        self.viscousity = float(viscousity)
        self.has_perlin_noise = bool(has_perlin_noise)
        self.fluid_color = fluid_color
        self.world_color = world_color

    @classmethod
    def default(cls):
        return cls()


class Rectangle:
    def __init__(self, down_left_point, up_right_point, fluid_container_size):
        self.down_left_point = tuple(int(v) for v in down_left_point)
        self.up_right_point = tuple(int(v) for v in up_right_point)
        self._approximate_points = [self.down_left_point, self.up_right_point]
        if not self.are_all_points_valid(int(fluid_container_size)):
            raise ValueError("Invalid input for Rectangle")

    @classmethod
    def default(cls):
        return cls((80, 80), (110, 110), SimulationConfigs.default().size)

    def are_all_points_valid(self, fluid_container_size):
        (x0, y0), (x1, y1) = self.down_left_point, self.up_right_point
        return _abi.lib().eqf_rect_validate(x0, y0, x1, y1, fluid_container_size) == _abi.EQF_OK

    def get_approximate_points(self):
        return self._approximate_points


class ObstaclesType:
    def __init__(self, rectangle):
        self.rectangle = rectangle

    @classmethod
    def Rectangle(cls, rectangle):
        return cls(rectangle)

    def get_approximate_points(self):
        return self.rectangle.get_approximate_points()


class GsMode(enum.IntEnum):
    Wavefront = _abi.GS_WAVEFRONT
    RedBlack = _abi.GS_REDBLACK


def _params(fluid_configs, simulation_configs, gs_mode, init_times, gs_iters=-1, pressure_iters=-1):
    p = _abi.default_params()
    p.size = simulation_configs.size
    p.delta_t = simulation_configs.delta_t
    p.frames = simulation_configs.frames
    p.diffusion = fluid_configs.diffusion
    p.viscosity = fluid_configs.viscousity
    p.gs_iters = gs_iters
    p.pressure_iters = pressure_iters
    p.gs_mode = int(gs_mode)
    p.init_times = init_times
    return p


class Fluid:
    """Device-resident fluid state with the reference's `Fluid` API."""

    def __init__(self, fluid_configs, simulation_configs, gs_mode=GsMode.Wavefront, *,
                 gs_iters=-1, pressure_iters=-1, _init_times=1, _scene=None):
        self.fluid_configs = fluid_configs
        self.simulation_configs = simulation_configs
        self.gs_mode = GsMode(gs_mode)
        if _scene is None:
            _scene = _abi.Scene(_params(fluid_configs, simulation_configs, gs_mode, _init_times,
                                        gs_iters, pressure_iters))
        self._scene = _scene

    # fluid.rs:93 / fluid.rs:83
    @classmethod
    def new(cls, init_fluid, init_simulation, gs_mode=GsMode.Wavefront, **kw):
        return cls(init_fluid, init_simulation, gs_mode, **kw)

    @classmethod
    def default(cls, gs_mode=GsMode.Wavefront):
        return cls(FluidConfigs.default(), SimulationConfigs.default(), gs_mode, _init_times=2)

    def clone(self):
        return Fluid(self.fluid_configs, self.simulation_configs, self.gs_mode,
                     _scene=self._scene.clone())

    @property
    def scene(self):
        return self._scene

    # fluid.rs:112-117
    @staticmethod
    def to_coordinate(idx, size):
        return (idx % size, idx // size)

    # pub fields (fluid.rs:61-80), read from the device
    @property
    def density(self):
        return self._scene.download(_abi.DENSITY)

    @property
    def velocities_x(self):
        return self._scene.download(_abi.VX)

    @property
    def velocities_y(self):
        return self._scene.download(_abi.VY)

    @property
    def cells_type(self):
        return [ContainerWall(int(v)) for v in self._scene.get_mask()]

    def field(self, f):
        return self._scene.download(f)

    # fluid.rs:120-131 (private in the reference, public here)
    def add_density(self, x, y, amount):
        self._scene.add_density(x, y, amount)

    def add_velocity(self, x, y, amount_x, amount_y):
        self._scene.add_velocity(x, y, amount_x, amount_y)

    # fluid.rs:437-524
    def step(self):
        self._sync_params()
        self._scene.step(1)

    def steps(self, n):
        self._sync_params()
        self._scene.step(n)

    def _sync_params(self):
        p = self._scene.params
        q = _params(self.fluid_configs, self.simulation_configs, self.gs_mode, p.init_times,
                    p.gs_iters, p.pressure_iters)
        if (q.delta_t, q.frames, q.diffusion, q.viscosity) != (p.delta_t, p.frames, p.diffusion,
                                                                p.viscosity):
            self._scene.set_params(q)

    # fluid.rs:610-619
    def fill_obstacle(self, obstacle):
        (x0, y0), (x1, y1) = obstacle.get_approximate_points()[:2]
        self._scene.fill_rect(x0, y0, x1, y1)

    # fluid.rs:575-599.  The reference draws an angle from noise-0.7 Perlin at
    # [dt, dt] and a point from rand's thread_rng, rotates the point about the
    # centre with geo-0.18 and adds twice its coordinates to the centre cell's
    # velocity.  Those crates are not available and thread_rng is unseeded, so
    # the angle (degrees, as geo's rotate) comes from the caller and the point
    # either from the caller or from `rng`, a seeded stand-in for thread_rng
    # (an object with randint(lo, hi), e.g. scenes.SplitMix64): parity
    # unpinned.  The rotation and the add run in the library (eqf_add_noise).
    def add_noise(self, angle_degrees, rand_x=None, rand_y=None, rng=None):
        size = self.simulation_configs.size
        if rand_x is None or rand_y is None:
            if rng is None:
                raise ValueError("add_noise needs the random point or a seeded rng")
            rand_x = rng.randint(0, size - 1)  # gen_range(0..size)
            rand_y = rng.randint(0, size - 1)
        self._scene.add_noise(angle_degrees, int(rand_x), int(rand_y))


class RenderingListener:
    """renderer_helpers.rs:87-165 (obstacles_color defaults to egui's
    Color32::RED)."""

    def __init__(self, obstacles_color=(255, 0, 0, 255)):
        self.obstacles_color = tuple(obstacles_color)

    def colors(self, fluid):
        fc = fluid.fluid_configs
        return _abi.make_colors(fc.fluid_color, fc.world_color, self.obstacles_color)

    def render_image(self, fluid):
        """Pixel (x, y) at [y, x] of a (size, size, 4) uint8 array."""
        size = fluid.simulation_configs.size
        return fluid.scene.render_rgba(self.colors(fluid)).reshape(size, size, 4)


class CurrentSimulation:
    """renderer_helpers.rs:29-81: obstacles are marked once, then each frame
    runs (noise,) step and hands the state to the consumer."""

    def __init__(self, fluid=None, obstacles=None):
        self.fluid = fluid if fluid is not None else Fluid.default()
        self.obstacles = obstacles if obstacles is not None else [
            ObstaclesType.Rectangle(Rectangle.default())]

    def mark_fluid_obstacles(self):
        for obstacle in self.obstacles:
            self.fluid.fill_obstacle(obstacle)

    def simulate(self, on_frame=None, noise=None):
        self.mark_fluid_obstacles()
        for i in range(self.fluid.simulation_configs.frames):
            if self.fluid.fluid_configs.has_perlin_noise and noise is not None:
                self.fluid.add_noise(*noise(i))
            self.fluid.step()
            if on_frame is not None:
                on_frame(i, self.fluid)

    def simulate_snapshots(self, noise=None, depth=2, listener=None):
        """Like ``simulate``, yielding (frame, density, rgba) per frame; the
        copy of frame i to the host runs behind the steps of frames i+1..,
        up to `depth` frames ahead.  rgba is the (size, size, 4) image of
        ``listener`` (None for density only)."""
        self.mark_fluid_obstacles()
        f = self.fluid
        f._sync_params()
        size = f.simulation_configs.size
        colors = None if listener is None else listener.colors(f)
        snap = _abi.Snapshot(f.scene, depth, listener is not None, colors)
        try:
            frames = f.simulation_configs.frames
            for i in range(frames):
                if i >= depth:  # every slot holds a capture: hand the oldest over
                    yield self._take(snap, size)
                if f.fluid_configs.has_perlin_noise and noise is not None:
                    f.add_noise(*noise(i))
                f.scene.step(1)
                snap.capture(i)
            for _ in range(min(depth, frames)):
                yield self._take(snap, size)
        finally:
            snap.close()

    @staticmethod
    def _take(snap, size):
        fr, dens, rgba = snap.acquire()
        snap.release()
        return fr, dens, None if rgba is None else rgba.reshape(size, size, 4)
