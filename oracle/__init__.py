"""ctypes front end of the CPU oracle (liboracle.so, built from eqf_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker, and by bench.py's cpu_baseline leg.  The product package
(equilibrium_amd) never imports it.

All arrays are numpy, row-major ``idx(x, y) = x + y*size`` (fluid.rs:31-35),
float32 for the *_f32 functions and float64 for the *_f64 ones; the mask is
uint8 with 0 = ContainerWall::DefaultWall and 1 = ContainerWall::NoWall.
"""
import contextlib
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

WALL = 0
FLUID = 1
ROW, COL, PASSIVE = 0, 1, 2
LEX, RB = 0, 1

_lib = None


def build():
    """Compile liboracle.so from eqf_oracle.c with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def load_library(path):
    """ctypes handle on a build of eqf_oracle.c at `path`, with every entry
    point's argument types declared."""
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    I64 = ctypes.c_int64
    for sfx, F in (("_f32", ctypes.c_float), ("_f64", ctypes.c_double)):
        getattr(L, "ora_set_bnd" + sfx).argtypes = [ctypes.c_int, P, I64, P]
        getattr(L, "ora_lin_solve" + sfx).argtypes = [ctypes.c_int, P, P, F, F, I64, I64, P, ctypes.c_int]
        getattr(L, "ora_diffuse" + sfx).argtypes = [ctypes.c_int, P, P, F, I64, F, I64, P, ctypes.c_int]
        getattr(L, "ora_project" + sfx).argtypes = [P, P, P, P, I64, I64, P, ctypes.c_int]
        getattr(L, "ora_advect" + sfx).argtypes = [ctypes.c_int, P, P, P, P, I64, F, P]
        getattr(L, "ora_step" + sfx).argtypes = [P] * 6 + [I64, F, F, F, I64, I64, P, ctypes.c_int]
        getattr(L, "ora_add_sources" + sfx).argtypes = [P] * 7 + [I64]
        getattr(L, "ora_run" + sfx).argtypes = [P] * 6 + [I64, F, F, F, I64, I64, P, ctypes.c_int, P, P, P, I64]
    L.ora_init_f32.argtypes = [P] * 7 + [ctypes.c_uint32, ctypes.c_int]
    L.ora_fill_rect.argtypes = [P, ctypes.c_uint32, I64, I64, I64, I64]
    L.ora_rect_valid.argtypes = [I64] * 5
    L.ora_rect_valid.restype = ctypes.c_int
    L.ora_to_coordinate.argtypes = [I64, I64, P, P]
    L.ora_idx.argtypes = [I64, I64, I64]
    L.ora_idx.restype = I64
    L.ora_step_par_f64.argtypes = [P] * 6 + [I64, ctypes.c_double, ctypes.c_double, ctypes.c_double, I64, I64, P]
    L.ora_step_par_f32.argtypes = [P] * 6 + [I64, ctypes.c_float, ctypes.c_float, ctypes.c_float, I64, I64, P]
    L.ora_poisson_residual_f64.argtypes = [P, P, I64, P, P]
    L.ora_poisson_residual_f64.restype = ctypes.c_double
    L.ora_render_rgba.argtypes = [P, P, I64, P, P, P, P]
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing; run `make -C oracle`")
        _lib = load_library(LIB_PATH)
    return _lib


@contextlib.contextmanager
def using(handle):
    """Routes every function of this module to another build of the oracle
    (a handle from load_library) inside the `with` block.  Used by the
    mutation tests, which check that planted mistakes fail the known-answer
    tests."""
    global _lib
    saved = lib()
    _lib = handle
    try:
        yield
    finally:
        _lib = saved


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def _sfx(a):
    if a.dtype == np.float32:
        return "_f32"
    if a.dtype == np.float64:
        return "_f64"
    raise TypeError(a.dtype)


FIELDS = ("density", "scratch", "vx", "vy", "vx0", "vy0")


class State:
    """The six fields and the mask of a `Fluid` (fluid.rs:51-81)."""

    def __init__(self, size, dtype=np.float32):
        n = size * size
        self.size = size
        self.f = {k: np.zeros(n, dtype) for k in FIELDS}
        self.mask = np.ones(n, np.uint8)

    @classmethod
    def new(cls, size, times=1):
        """`Fluid::new` (times=1) or `Fluid::default` (times=2) state."""
        st = cls(size, np.float32)
        f = st.f
        lib().ora_init_f32(*(_p(f[k]) for k in FIELDS), _p(st.mask), size, times)
        return st

    def astype(self, dtype):
        st = State(self.size, dtype)
        for k in FIELDS:
            st.f[k] = self.f[k].astype(dtype)
        st.mask = self.mask.copy()
        return st

    def copy(self):
        st = State(self.size, self.f["density"].dtype)
        for k in FIELDS:
            st.f[k] = self.f[k].copy()
        st.mask = self.mask.copy()
        return st

    def fill_rect(self, x0, y0, x1, y1):
        lib().ora_fill_rect(_p(self.mask), self.size, x0, y0, x1, y1)

    def run(self, dt, diffusion, viscosity, iters, nframes, press_iters=None, order=LEX,
            src_d=None, src_u=None, src_v=None):
        f = self.f
        dtype = f["density"].dtype
        sfx = _sfx(f["density"])
        conv = lambda a: None if a is None else np.ascontiguousarray(a, dtype)
        sd, su, sv = conv(src_d), conv(src_u), conv(src_v)
        getattr(lib(), "ora_run" + sfx)(
            *(_p(f[k]) for k in FIELDS), self.size, dt, diffusion, viscosity, iters,
            iters if press_iters is None else press_iters, _p(self.mask), order,
            _p(sd), _p(su), _p(sv), nframes)

    def step_par(self, dt, diffusion, viscosity, iters, press_iters=None):
        """One red-black step (OpenMP), same bits as run(..., order=RB) for
        one frame without sources; f32 or f64 by the state's dtype."""
        f = self.f
        getattr(lib(), "ora_step_par" + _sfx(f["density"]))(
            *(_p(f[k]) for k in FIELDS), self.size, dt, diffusion, viscosity, iters,
            iters if press_iters is None else press_iters, _p(self.mask))

    def step_par_f64(self, dt, diffusion, viscosity, iters, press_iters=None):
        assert self.f["density"].dtype == np.float64
        self.step_par(dt, diffusion, viscosity, iters, press_iters)


def set_bnd(mode, x, size, mask):
    getattr(lib(), "ora_set_bnd" + _sfx(x))(mode, _p(x), size, _p(mask))


def lin_solve(mode, x, x0, a, c, size, iters, mask, order=LEX):
    getattr(lib(), "ora_lin_solve" + _sfx(x))(mode, _p(x), _p(x0), a, c, size, iters, _p(mask), order)


def diffuse(mode, x, x0, k, size, dt, iters, mask, order=LEX):
    getattr(lib(), "ora_diffuse" + _sfx(x))(mode, _p(x), _p(x0), k, size, dt, iters, _p(mask), order)


def project(u, v, p, div, size, iters, mask, order=LEX):
    getattr(lib(), "ora_project" + _sfx(u))(_p(u), _p(v), _p(p), _p(div), size, iters, _p(mask), order)


def advect(mode, d, d0, u, v, size, dt, mask):
    getattr(lib(), "ora_advect" + _sfx(d))(mode, _p(d), _p(d0), _p(u), _p(v), size, dt, _p(mask))


def rect_valid(x0, y0, x1, y1, size):
    return bool(lib().ora_rect_valid(x0, y0, x1, y1, size))


def to_coordinate(idx, size):
    x = ctypes.c_int64()
    y = ctypes.c_int64()
    lib().ora_to_coordinate(idx, size, ctypes.byref(x), ctypes.byref(y))
    return x.value, y.value


def idx(x, y, size):
    return lib().ora_idx(x, y, size)


def poisson_residual(p, div, size):
    rr = ctypes.c_double()
    rd = ctypes.c_double()
    mx = lib().ora_poisson_residual_f64(_p(np.ascontiguousarray(p, np.float64)),
                                        _p(np.ascontiguousarray(div, np.float64)), size,
                                        ctypes.byref(rr), ctypes.byref(rd))
    return mx, rr.value, rd.value


FLUID_COLOR = (208, 88, 157, 220)     # configs.rs:56
WORLD_COLOR = (94, 146, 162, 128)     # configs.rs:57
OBSTACLE_COLOR = (255, 0, 0, 255)     # egui Color32::RED (renderer_helpers.rs:101)


def render_rgba(density, mask, size, fluid=FLUID_COLOR, world=WORLD_COLOR, obstacle=OBSTACLE_COLOR):
    """renderer_helpers.rs:143-165: (size*size, 4) uint8 RGBA image."""
    d = np.ascontiguousarray(density, np.float32).ravel()
    m = np.ascontiguousarray(mask, np.uint8).ravel()
    assert d.size == size * size and m.size == size * size
    out = np.empty((size * size, 4), np.uint8)
    cols = [np.ascontiguousarray(np.array(c, np.uint8)) for c in (fluid, world, obstacle)]
    lib().ora_render_rgba(_p(d), _p(m), size, *(_p(c) for c in cols), _p(out))
    return out
