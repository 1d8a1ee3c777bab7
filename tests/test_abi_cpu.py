"""CPU-side checks of the C ABI and the host mirror (no GPU needed)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from equilibrium_amd import _abi, scenes
from equilibrium_amd.fluid import FluidConfigs, Rectangle, SimulationConfigs


def test_library_exports_every_declared_symbol():
    decl = _abi.declared_functions()
    assert len(decl) >= 25
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [f for f in decl if f not in exported]
    assert not missing, missing
    L = _abi.lib()
    for f in decl:
        assert getattr(L, f) is not None


def test_params_default_mirror_reference_defaults():
    """SimulationConfigs::default (configs.rs:15-21), FluidConfigs::default (configs.rs:51-59)."""
    p = _abi.default_params()
    assert (p.size, p.frames) == (128, 16)
    assert p.delta_t == pytest.approx(0.02) and p.viscosity == pytest.approx(0.001)
    assert p.diffusion == 0.0 and p.gs_iters == -1 and p.gs_mode == _abi.GS_WAVEFRONT
    s, f = SimulationConfigs.default(), FluidConfigs.default()
    assert (s.delta_t, s.frames, s.size) == (0.02, 16, 128)
    assert (f.diffusion, f.viscousity, f.has_perlin_noise) == (0.0, 0.001, True)


def test_rect_validate_and_error_message():
    L = _abi.lib()
    assert L.eqf_rect_validate(80, 80, 110, 110, 128) == _abi.EQF_OK
    assert L.eqf_rect_validate(50, 120, 127, 110, 128) == _abi.EQF_EINVAL
    assert _abi.last_error() == "Invalid input for Rectangle"
    assert L.eqf_rect_validate(12, 12, 10, 10, 128) == _abi.EQF_EINVAL


def test_python_rectangle_raises_like_reference_panics():
    """obstacle.rs:100-115 should_panic cases raise with the reference's message."""
    with pytest.raises(ValueError, match="Invalid input for Rectangle"):
        Rectangle((50, 120), (127, 110), 128)
    with pytest.raises(ValueError, match="Invalid input for Rectangle"):
        Rectangle((12, 12), (10, 10), 128)
    r = Rectangle.default()
    assert r.get_approximate_points() == [(80, 80), (110, 110)]


def test_null_and_invalid_arguments_fail_cleanly():
    L = _abi.lib()
    p = _abi.default_params()
    h = ctypes.c_void_p()
    assert L.eqf_scene_create(None, ctypes.byref(p), ctypes.byref(h)) == _abi.EQF_EINVAL
    assert L.eqf_step(None, 1) == _abi.EQF_EINVAL
    assert L.eqf_download(None, 0, None) == _abi.EQF_EINVAL
    assert L.eqf_batch_step(None, 1) == _abi.EQF_EINVAL
    assert L.eqf_batch_set_schedule(None, _abi.SCHED_RING) == _abi.EQF_EINVAL
    assert "NULL" in _abi.last_error() or "null" in _abi.last_error().lower()


def test_ctx_create_without_gpu_reports_device_error():
    h = ctypes.c_void_p()
    rc = _abi.lib().eqf_ctx_create(0, ctypes.byref(h))
    if rc == _abi.EQF_OK:  # a GPU is present: nothing to check here
        _abi.lib().eqf_ctx_destroy(h)
        pytest.skip("a GPU is present")
    assert rc in (_abi.EQF_EDEVICE, _abi.EQF_EINVAL)
    assert _abi.last_error()


def test_scene_generator_is_deterministic():
    a = scenes.batched_mix(n=6, frames=5)
    b = scenes.batched_mix(n=6, frames=5)
    for x, y in zip(a, b):
        assert x["size"] == y["size"] and x["rects"] == y["rects"]
        assert np.array_equal(x["src_d"], y["src_d"]) and np.array_equal(x["src_u"], y["src_u"])
    sizes = {sc["size"] for sc in scenes.batched_mix(n=64, frames=1)}
    assert sizes <= {66, 130, 258, 514} and len(sizes) == 4


def test_rectangles_do_not_overlap_and_stay_inside():
    for i in range(20):
        sc = scenes.stam_scene(128, scenes.derive_seed(900, i), n_rects=4)
        for r in sc["rects"]:
            assert 1 <= r[0] < r[2] <= 129 and 1 <= r[1] < r[3] <= 129
        for a_ in range(len(sc["rects"])):
            for b_ in range(a_ + 1, len(sc["rects"])):
                assert not scenes._overlap(sc["rects"][a_], sc["rects"][b_])


def test_default_colours_mirror_reference():
    """FluidConfigs::default colours (configs.rs:56-57) and the renderer's
    obstacle colour, egui's Color32::RED (renderer_helpers.rs:101)."""
    c = _abi.make_colors()
    assert list(c.fluid) == [208, 88, 157, 220]
    assert list(c.world) == [94, 146, 162, 128]
    assert list(c.obstacle) == [255, 0, 0, 255]
    fc = FluidConfigs.default()
    assert tuple(fc.fluid_color) == tuple(c.fluid) and tuple(fc.world_color) == tuple(c.world)
    c2 = _abi.make_colors(fluid=(1, 2, 3, 4))
    assert list(c2.fluid) == [1, 2, 3, 4] and list(c2.world) == [94, 146, 162, 128]


def test_render_and_snapshot_reject_null_arguments():
    L = _abi.lib()
    assert L.eqf_render_rgba(None, None, None) == _abi.EQF_EINVAL
    assert L.eqf_snapshot_create(None, 2, 0, None, None) == _abi.EQF_EINVAL
    assert L.eqf_snapshot_capture(None, 0) == _abi.EQF_EINVAL
    assert L.eqf_snapshot_acquire(None, None, None, None) == _abi.EQF_EINVAL
    assert L.eqf_snapshot_release(None) == _abi.EQF_EINVAL
    assert L.eqf_snapshot_destroy(None) == _abi.EQF_OK


def test_batch_rejects_a_scene_listed_twice():
    """A scene appearing twice in one batch would be advanced twice per frame
    by concurrent workgroups; eqf_batch_create refuses it before touching the
    device or the scenes (the handles below are never dereferenced)."""
    L = _abi.lib()
    fake_ctx = ctypes.create_string_buffer(64)
    fake = ctypes.create_string_buffer(64)
    other = ctypes.create_string_buffer(64)
    arr = (ctypes.c_void_p * 3)(ctypes.addressof(fake), ctypes.addressof(other), ctypes.addressof(fake))
    h = ctypes.c_void_p()
    assert L.eqf_batch_create(ctypes.addressof(fake_ctx), arr, 3, ctypes.byref(h)) == _abi.EQF_EINVAL
    assert "more than once" in _abi.last_error()


def test_scene_parameter_limits_rejected_before_the_device():
    """eqf_scene_create validates the parameters before any device call (the
    context below is never dereferenced): size >= 3 (one interior cell inside
    the wall ring, fluid.rs:552-570), wavefront mode up to size 1026 (16 strips
    of 64 rows in one workgroup), a known GS mode, init run once or twice
    (Fluid::new / Fluid::default, fluid.rs:83-110), bounded iteration counts."""
    L = _abi.lib()
    fake_ctx = ctypes.create_string_buffer(64)
    h = ctypes.c_void_p()

    def create(**kw):
        p = _abi.default_params()
        for k, v in kw.items():
            setattr(p, k, v)
        return L.eqf_scene_create(ctypes.addressof(fake_ctx), ctypes.byref(p), ctypes.byref(h))

    cases = [({"size": 2}, "size must be >= 3"),
             ({"size": 1027, "gs_mode": _abi.GS_WAVEFRONT}, "size <= 1026"),
             ({"size": 40000}, "size must be <= 32768"),
             ({"gs_mode": 7}, "gs_mode"),
             ({"init_times": 3}, "init_times"),
             ({"gs_iters": 1 << 25}, "iteration count too large")]
    for kw, msg in cases:
        assert create(**kw) == _abi.EQF_EINVAL, kw
        assert msg in _abi.last_error(), (kw, _abi.last_error())
    assert not h.value
