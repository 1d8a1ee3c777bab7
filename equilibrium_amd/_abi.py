"""ctypes binding of libeqfluid.so (include/eqfluid.h).

Loads the in-tree library built by ``equilibrium_amd.build`` and raises if it
is missing: there is no CPU fallback for any entry point.
"""
import ctypes
import os
import re

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "libeqfluid.so")
HEADER_PATH = os.path.join(REPO_DIR, "include", "eqfluid.h")

EQF_OK, EQF_EINVAL, EQF_ENOMEM, EQF_EDEVICE, EQF_EBUSY, EQF_ENOTFOUND = range(6)
DENSITY, SCRATCH, VX, VY, VX0, VY0 = range(6)
DEFAULT_WALL, NO_WALL = 0, 1
GS_WAVEFRONT, GS_REDBLACK = 0, 1
BATCH_LARGEST_FIRST, BATCH_GIVEN_ORDER = 0, 1
SCHED_AUTO, SCHED_WIDE, SCHED_RING, SCHED_FINE = 0, 1, 2, 3


class EqfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"eqf error {code}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [
        ("size", ctypes.c_uint32),
        ("delta_t", ctypes.c_float),
        ("frames", ctypes.c_int64),
        ("diffusion", ctypes.c_float),
        ("viscosity", ctypes.c_float),
        ("gs_iters", ctypes.c_int64),
        ("pressure_iters", ctypes.c_int64),
        ("gs_mode", ctypes.c_int32),
        ("init_times", ctypes.c_int32),
    ]


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("max_active", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class Request(ctypes.Structure):
    _fields_ = [("params", Params), ("nframes", ctypes.c_int64), ("n_rects", ctypes.c_int32),
                ("keep_velocity", ctypes.c_int32), ("borrow_inputs", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("rects", ctypes.c_void_p),
                ("src_density", ctypes.c_void_p), ("src_vx", ctypes.c_void_p),
                ("src_vy", ctypes.c_void_p), ("init_vx", ctypes.c_void_p), ("init_vy", ctypes.c_void_p)]


class RequestStatus(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int32), ("reserved", ctypes.c_int32), ("frames_done", ctypes.c_int64),
                ("submit_ms", ctypes.c_double), ("admit_ms", ctypes.c_double), ("done_ms", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [("ticks", ctypes.c_int64), ("scene_frames", ctypes.c_int64),
                ("cell_updates", ctypes.c_double), ("busy_ms", ctypes.c_double),
                ("p50_frame_ms", ctypes.c_double), ("p99_frame_ms", ctypes.c_double),
                ("p50_queue_ms", ctypes.c_double), ("p99_queue_ms", ctypes.c_double),
                ("max_resident", ctypes.c_int64), ("peak_bytes", ctypes.c_uint64)]


REQ_QUEUED, REQ_ACTIVE, REQ_DONE, REQ_FAILED, REQ_CANCELLED = range(5)


class Colors(ctypes.Structure):
    _fields_ = [("fluid", ctypes.c_uint8 * 4), ("world", ctypes.c_uint8 * 4), ("obstacle", ctypes.c_uint8 * 4)]


def make_colors(fluid=None, world=None, obstacle=None):
    """eqf_colors with the library defaults where an argument is None."""
    c = Colors()
    lib().eqf_colors_default(ctypes.byref(c))
    for name, val in (("fluid", fluid), ("world", world), ("obstacle", obstacle)):
        if val is not None:
            getattr(c, name)[:] = [int(v) for v in val]
    return c

_lib = None


def declared_functions():
    """Names of every function declared in include/eqfluid.h."""
    text = open(HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(eqf_[a-z0-9_]+)\s*\(", text)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build the HIP library first "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        PP = ctypes.POINTER(ctypes.c_void_p)
        I = ctypes.c_int
        I64 = ctypes.c_int64
        U32 = ctypes.c_uint32
        F = ctypes.c_float
        sig = {
            "eqf_params_default": (None, [ctypes.POINTER(Params)]),
            "eqf_last_error": (ctypes.c_char_p, []),
            "eqf_ctx_create": (I, [I, PP]),
            "eqf_ctx_destroy": (I, [P]),
            "eqf_ctx_device_key": (I, [P, ctypes.c_char_p, ctypes.c_size_t]),
            "eqf_scene_create": (I, [P, ctypes.POINTER(Params), PP]),
            "eqf_scene_clone": (I, [P, PP]),
            "eqf_scene_destroy": (I, [P]),
            "eqf_scene_get_params": (I, [P, ctypes.POINTER(Params)]),
            "eqf_scene_set_params": (I, [P, ctypes.POINTER(Params)]),
            "eqf_rect_validate": (I, [I64, I64, I64, I64, U32]),
            "eqf_fill_rect": (I, [P, I64, I64, I64, I64]),
            "eqf_set_mask": (I, [P, P]),
            "eqf_get_mask": (I, [P, P]),
            "eqf_add_density": (I, [P, U32, U32, F]),
            "eqf_add_velocity": (I, [P, U32, U32, F, F]),
            "eqf_add_cell": (I, [P, I, U32, U32, F]),
            "eqf_add_noise": (I, [P, F, U32, U32]),
            "eqf_add_source": (I, [P, I, P]),
            "eqf_set_frame_source": (I, [P, I, P]),
            "eqf_upload": (I, [P, I, P]),
            "eqf_download": (I, [P, I, P]),
            "eqf_step": (I, [P, I64]),
            "eqf_sync": (I, [P]),
            "eqf_scene_set_tuning": (I, [P, I, I]),
            "eqf_batch_create": (I, [P, PP, I, PP]),
            "eqf_batch_step": (I, [P, I64]),
            "eqf_batch_sync": (I, [P]),
            "eqf_batch_destroy": (I, [P]),
            "eqf_batch_last_ms": (I, [P, ctypes.POINTER(ctypes.c_double)]),
            "eqf_batch_set_tuning": (I, [P, I, I]),
            "eqf_batch_set_schedule": (I, [P, I]),
            "eqf_batch_set_band": (I, [P, I, I]),
            "eqf_server_create": (I, [P, ctypes.POINTER(ServerConfig), PP]),
            "eqf_server_submit": (I, [P, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]),
            "eqf_server_poll": (I, [P, ctypes.c_uint64, ctypes.POINTER(RequestStatus)]),
            "eqf_server_result": (I, [P, ctypes.c_uint64, I, P]),
            "eqf_server_release": (I, [P, ctypes.c_uint64]),
            "eqf_server_wait_idle": (I, [P, ctypes.c_double]),
            "eqf_server_stats_get": (I, [P, ctypes.POINTER(ServerStats)]),
            "eqf_server_stats_sizes": (I, [P, U32, U32, ctypes.POINTER(ServerStats)]),
            "eqf_server_destroy": (I, [P]),
            "eqf_colors_default": (None, [ctypes.POINTER(Colors)]),
            "eqf_render_rgba": (I, [P, ctypes.POINTER(Colors), P]),
            "eqf_render_rgba_noalloc": (I, [P, ctypes.POINTER(Colors), P]),
            "eqf_snapshot_create": (I, [P, I, I, ctypes.POINTER(Colors), PP]),
            "eqf_snapshot_capture": (I, [P, I64]),
            "eqf_snapshot_acquire": (I, [P, ctypes.POINTER(I64), ctypes.POINTER(ctypes.POINTER(F)),
                                         ctypes.POINTER(ctypes.POINTER(ctypes.c_uint8))]),
            "eqf_snapshot_release": (I, [P]),
            "eqf_snapshot_destroy": (I, [P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error():
    return lib().eqf_last_error().decode()


def check(rc):
    if rc != EQF_OK:
        raise EqfError(rc, last_error())
    return rc


def ptr(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def default_params():
    p = Params()
    lib().eqf_params_default(ctypes.byref(p))
    return p


class Context:
    def __init__(self, device=0):
        h = ctypes.c_void_p()
        check(lib().eqf_ctx_create(device, ctypes.byref(h)))
        self.h = h

    def device_key(self):
        """'<device name>/<gfx arch>/<CU count>' of the context's GPU."""
        buf = ctypes.create_string_buffer(512)
        check(lib().eqf_ctx_device_key(self.h, buf, len(buf)))
        return buf.value.decode(errors="replace")

    def close(self):
        if self.h:
            lib().eqf_ctx_destroy(self.h)
            self.h = None


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class Scene:
    """One device-resident `Fluid` state (thin handle over eqf_scene)."""

    def __init__(self, params, ctx=None, _handle=None):
        self.ctx = ctx or default_context()
        if _handle is not None:
            self.h = _handle
        else:
            h = ctypes.c_void_p()
            check(lib().eqf_scene_create(self.ctx.h, ctypes.byref(params), ctypes.byref(h)))
            self.h = h
        p = Params()
        check(lib().eqf_scene_get_params(self.h, ctypes.byref(p)))
        self.params = p
        self.size = p.size

    def clone(self):
        h = ctypes.c_void_p()
        check(lib().eqf_scene_clone(self.h, ctypes.byref(h)))
        return Scene(None, self.ctx, _handle=h)

    def close(self):
        if getattr(self, "h", None):
            check(lib().eqf_scene_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        check(lib().eqf_scene_set_params(self.h, ctypes.byref(p)))
        self.params = p

    def fill_rect(self, x0, y0, x1, y1):
        check(lib().eqf_fill_rect(self.h, x0, y0, x1, y1))

    def set_mask(self, mask):
        m = np.ascontiguousarray(mask, np.uint8)
        check(lib().eqf_set_mask(self.h, ptr(m)))

    def get_mask(self):
        m = np.empty(self.size * self.size, np.uint8)
        check(lib().eqf_get_mask(self.h, ptr(m)))
        return m

    def add_density(self, x, y, amount):
        check(lib().eqf_add_density(self.h, x, y, amount))

    def add_velocity(self, x, y, ax, ay):
        check(lib().eqf_add_velocity(self.h, x, y, ax, ay))

    def add_cell(self, field, x, y, amount):
        check(lib().eqf_add_cell(self.h, field, x, y, amount))

    def add_noise(self, angle_degrees, rand_x, rand_y):
        check(lib().eqf_add_noise(self.h, angle_degrees, rand_x, rand_y))

    def add_source(self, field, amounts):
        a = np.ascontiguousarray(amounts, np.float32).ravel()
        check(lib().eqf_add_source(self.h, field, ptr(a)))

    def set_frame_source(self, field, amounts):
        a = None if amounts is None else np.ascontiguousarray(amounts, np.float32).ravel()
        check(lib().eqf_set_frame_source(self.h, field, ptr(a)))

    def upload(self, field, arr):
        a = np.ascontiguousarray(arr, np.float32).ravel()
        assert a.size == self.size * self.size
        check(lib().eqf_upload(self.h, field, ptr(a)))

    def download(self, field):
        a = np.empty(self.size * self.size, np.float32)
        check(lib().eqf_download(self.h, field, ptr(a)))
        return a

    def step(self, nframes=1):
        check(lib().eqf_step(self.h, nframes))

    def sync(self):
        check(lib().eqf_sync(self.h))

    def set_tuning(self, tile, sweeps_per_block):
        check(lib().eqf_scene_set_tuning(self.h, tile, sweeps_per_block))

    def render_rgba(self, colors=None):
        """(size*size, 4) uint8 image of the current state (eqf_render_rgba)."""
        out = np.empty((self.size * self.size, 4), np.uint8)
        check(lib().eqf_render_rgba(self.h, None if colors is None else ctypes.byref(colors), ptr(out)))
        return out

    def render_rgba_noalloc(self, colors=None, out=None):
        """Same image through eqf_render_rgba_noalloc (no device allocation;
        fixed per-GPU scratch); `out` may be a preallocated (size*size, 4)
        uint8 array."""
        if out is None:
            out = np.empty((self.size * self.size, 4), np.uint8)
        assert out.dtype == np.uint8 and out.size == 4 * self.size * self.size and out.flags["C_CONTIGUOUS"]
        check(lib().eqf_render_rgba_noalloc(self.h, None if colors is None else ctypes.byref(colors), ptr(out)))
        return out


class Snapshot:
    """Ring of per-frame host snapshots of one scene (eqf_snapshot_*)."""

    def __init__(self, scene, depth=2, with_rgba=False, colors=None):
        self.scene = scene
        self.size = scene.size
        self.with_rgba = bool(with_rgba)
        h = ctypes.c_void_p()
        check(lib().eqf_snapshot_create(scene.h, depth, 1 if with_rgba else 0,
                                        None if colors is None else ctypes.byref(colors), ctypes.byref(h)))
        self.h = h

    def capture(self, frame):
        check(lib().eqf_snapshot_capture(self.h, frame))

    def acquire(self, copy=True):
        """(frame, density, rgba or None) of the oldest capture.  With
        copy=False the arrays view the pinned slot and stay valid only until
        release()."""
        fr = ctypes.c_int64()
        d = ctypes.POINTER(ctypes.c_float)()
        c = ctypes.POINTER(ctypes.c_uint8)()
        check(lib().eqf_snapshot_acquire(self.h, ctypes.byref(fr), ctypes.byref(d), ctypes.byref(c)))
        n = self.size * self.size
        dens = np.ctypeslib.as_array(d, shape=(n,))
        rgba = np.ctypeslib.as_array(c, shape=(n, 4)) if self.with_rgba else None
        if copy:
            dens = dens.copy()
            rgba = None if rgba is None else rgba.copy()
        return fr.value, dens, rgba

    def release(self):
        check(lib().eqf_snapshot_release(self.h))

    def close(self):
        if getattr(self, "h", None):
            check(lib().eqf_snapshot_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    def __init__(self, scenes, ctx=None):
        self.ctx = ctx or scenes[0].ctx
        self.scenes = list(scenes)
        arr = (ctypes.c_void_p * len(scenes))(*[s.h for s in scenes])
        h = ctypes.c_void_p()
        check(lib().eqf_batch_create(self.ctx.h, arr, len(scenes), ctypes.byref(h)))
        self.h = h

    def step(self, nframes=1):
        check(lib().eqf_batch_step(self.h, nframes))

    def sync(self):
        check(lib().eqf_batch_sync(self.h))

    def last_ms(self):
        ms = ctypes.c_double()
        check(lib().eqf_batch_last_ms(self.h, ctypes.byref(ms)))
        return ms.value

    def set_tuning(self, max_waves, order=BATCH_LARGEST_FIRST):
        check(lib().eqf_batch_set_tuning(self.h, max_waves, order))

    def set_schedule(self, sched):
        """Wavefront schedule of a wavefront-mode batch (SCHED_AUTO, SCHED_WIDE, SCHED_RING)."""
        check(lib().eqf_batch_set_schedule(self.h, sched))

    def set_band(self, tile, sweeps_per_block):
        """Band shape and sweeps per launch of a red-black batch."""
        check(lib().eqf_batch_set_band(self.h, tile, sweeps_per_block))

    def close(self):
        if getattr(self, "h", None):
            check(lib().eqf_batch_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching server (eqf_server_*)."""

    def __init__(self, hbm_budget_bytes, max_active=0, ctx=None):
        self.ctx = ctx or default_context()
        cfg = ServerConfig(hbm_budget_bytes, max_active, 0)
        h = ctypes.c_void_p()
        check(lib().eqf_server_create(self.ctx.h, ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self._keep = {}
        self._borrowed = {}

    def submit(self, sc, nframes=None, keep_velocity=False, gs_mode=GS_WAVEFRONT, borrow=False):
        """Submits a scene dict (equilibrium_amd.scenes format).  borrow=True
        lets the server read the scene's arrays in place at admission (this
        object keeps them referenced until close())."""
        from . import scenes as _sc
        p = _sc.scene_params(sc, gs_mode)
        S = sc["size"]
        arrs = {}
        for k in ("src_d", "src_u", "src_v", "init_u", "init_v"):
            a = sc.get(k)
            arrs[k] = None if a is None else np.ascontiguousarray(a, np.float32).ravel()
        rects = np.ascontiguousarray(np.array(sc["rects"], np.int64).reshape(-1, 4))
        req = Request(p, sc["frames"] if nframes is None else nframes, len(sc["rects"]),
                      1 if keep_velocity else 0, 1 if borrow else 0, 0,
                      rects.ctypes.data if len(sc["rects"]) else None,
                      *(None if arrs[k] is None else arrs[k].ctypes.data
                        for k in ("src_d", "src_u", "src_v", "init_u", "init_v")))
        t = ctypes.c_uint64()
        check(lib().eqf_server_submit(self.h, ctypes.byref(req), ctypes.byref(t)))
        self._keep[t.value] = S
        if borrow:
            self._borrowed[t.value] = (rects, arrs)
        return t.value

    def poll(self, ticket):
        st = RequestStatus()
        check(lib().eqf_server_poll(self.h, ticket, ctypes.byref(st)))
        return st

    def result(self, ticket, field=DENSITY):
        S = self._keep[ticket]
        out = np.empty(S * S, np.float32)
        check(lib().eqf_server_result(self.h, ticket, field, ptr(out)))
        return out

    def release(self, ticket):
        """Forgets a ticket (cancelling it if queued or resident); borrowed
        arrays stay referenced until close()."""
        check(lib().eqf_server_release(self.h, ticket))
        self._keep.pop(ticket, None)

    def wait_idle(self, timeout_ms=-1.0):
        check(lib().eqf_server_wait_idle(self.h, timeout_ms))

    def stats(self, min_size=None, max_size=None):
        """Server statistics; with min_size / max_size, latencies and counts
        of the requests of those sizes only (ring-inclusive sizes)."""
        st = ServerStats()
        if min_size is None and max_size is None:
            check(lib().eqf_server_stats_get(self.h, ctypes.byref(st)))
        else:
            check(lib().eqf_server_stats_sizes(self.h, 0 if min_size is None else min_size,
                                               0xFFFFFFFF if max_size is None else max_size, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "h", None):
            check(lib().eqf_server_destroy(self.h))
            self.h = None
        self._borrowed = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
