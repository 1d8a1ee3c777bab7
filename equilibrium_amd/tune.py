"""Autotuners.

* Red-black mode: band shape x sweeps per launch (eqf_scene_set_tuning).
* Batches (wavefront mode): waves per workgroup at most -- with fewer waves a
  strip runs fewer sweep streams -- and the scene-to-workgroup order
  (eqf_batch_set_tuning), and the wavefront schedule (eqf_batch_set_schedule).
* Batches (red-black mode): band shape x sweeps per launch over all scenes
  (eqf_batch_set_band).

Each times frames per configuration on the GPU (on clones, so the tuned
scenes are not advanced) and keeps the fastest; results are cached in a JSON
file (EQF_TUNE_CACHE, default equilibrium_amd/tuning_cache.json) keyed by the
workload and the device (name, gfx architecture and CU count, as the library
reports them for the scenes' context), so later runs reuse them.
"""
import hashlib
import json
import os
import time

from . import _abi

# band shape index -> (waves per workgroup, rows per wave); regions are 128 columns wide
TILES = {0: (8, 16), 1: (8, 32), 2: (16, 16), 3: (8, 18), 4: (8, 14), 5: (8, 10)}
SWEEPS_PER_BLOCK = (2, 4, 5, 6, 8)
CACHE = os.environ.get("EQF_TUNE_CACHE", os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                       "tuning_cache.json"))


def _device_key(ctx):
    return ctx.device_key()


def _load():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(db):
    try:
        with open(CACHE, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
    except OSError:
        pass


def tune_redblack(scene, frames=1, use_cache=True):
    """Picks (tile, sweeps_per_block) for a red-black _abi.Scene, applies it and
    returns (tile, sweeps, timings).  The scene is advanced by the trial
    frames; callers that need an untouched state tune a clone."""
    p = scene.params
    key = f"rbband:{p.size}:{p.gs_iters}:{p.pressure_iters}:{_device_key(scene.ctx)}"
    db = _load() if use_cache else {}
    if key in db:
        tile, tsw = db[key]["best"]
        scene.set_tuning(tile, tsw)
        return tile, tsw, db[key].get("ms", {})
    timings = {}
    for tile in TILES:
        for tsw in SWEEPS_PER_BLOCK:
            scene.set_tuning(tile, tsw)
            scene.step(1)  # builds the graph
            scene.sync()
            t0 = time.perf_counter()
            scene.step(frames)
            scene.sync()
            timings[f"{tile}:{tsw}"] = (time.perf_counter() - t0) * 1e3 / frames
    best = min(timings, key=timings.get)
    tile, tsw = (int(x) for x in best.split(":"))
    scene.set_tuning(tile, tsw)
    db[key] = {"best": [tile, tsw], "ms": timings, "tiles": {str(k): v for k, v in TILES.items()}}
    _save(db)
    return tile, tsw, timings


BATCH_WAVES = (16, 8, 4, 2, 1)
BATCH_ORDERS = (_abi.BATCH_LARGEST_FIRST, _abi.BATCH_GIVEN_ORDER)
BATCH_SCHEDS = (_abi.SCHED_WIDE, _abi.SCHED_RING, _abi.SCHED_FINE)


def _batch_key(scenes):
    sig = ",".join(f"{s.params.size}:{s.params.gs_iters}:{s.params.pressure_iters}" for s in scenes)
    return f"batch2:{len(scenes)}:{hashlib.sha1(sig.encode()).hexdigest()[:16]}:{_device_key(scenes[0].ctx)}"


def tune_batch(scenes, frames=2, use_cache=True):
    """Picks (max_waves, order, schedule) for a batch of the given
    wavefront-mode _abi.Scene objects and returns (max_waves, order,
    schedule, timings in ms per frame).  Trials run on clones."""
    key = _batch_key(scenes)
    db = _load() if use_cache else {}
    if key in db:
        w, o, sc = db[key]["best"]
        return w, o, sc, db[key].get("ms", {})
    maxns = max((s.params.size - 2 + 63) // 64 for s in scenes)
    clones = [s.clone() for s in scenes]
    b = _abi.Batch(clones)
    timings = {}
    try:
        for w in BATCH_WAVES:
            if w < maxns:
                continue
            for o in BATCH_ORDERS:
                b.set_tuning(w, o)
                for sc in BATCH_SCHEDS:
                    b.set_schedule(sc)
                    b.step(frames)  # builds the graph of `frames` launches
                    b.sync()
                    t0 = time.perf_counter()
                    b.step(frames)
                    b.sync()
                    timings[f"{w}:{o}:{sc}"] = (time.perf_counter() - t0) * 1e3 / frames
    finally:
        b.close()
        for c in clones:
            c.close()
    best = min(timings, key=timings.get)
    w, o, sc = (int(x) for x in best.split(":"))
    db = _load() if use_cache else {}
    db[key] = {"best": [w, o, sc], "ms": timings}
    if use_cache:
        _save(db)
    return w, o, sc, timings


def tune_batch_band(scenes, frames=2, use_cache=True):
    """Picks (tile, sweeps_per_block) for a batch of red-black _abi.Scene
    objects; returns (tile, sweeps, timings in ms per frame).  Trials run on
    clones."""
    key = "rb" + _batch_key(scenes)
    db = _load() if use_cache else {}
    if key in db:
        tile, tsw = db[key]["best"]
        return tile, tsw, db[key].get("ms", {})
    clones = [s.clone() for s in scenes]
    b = _abi.Batch(clones)
    timings = {}
    try:
        for tile in TILES:
            for tsw in SWEEPS_PER_BLOCK:
                b.set_band(tile, tsw)
                b.step(frames)  # builds the tables and the graph
                b.sync()
                t0 = time.perf_counter()
                b.step(frames)
                b.sync()
                timings[f"{tile}:{tsw}"] = (time.perf_counter() - t0) * 1e3 / frames
    finally:
        b.close()
        for c in clones:
            c.close()
    best = min(timings, key=timings.get)
    tile, tsw = (int(x) for x in best.split(":"))
    db = _load() if use_cache else {}
    db[key] = {"best": [tile, tsw], "ms": timings, "tiles": {str(k): v for k, v in TILES.items()}}
    if use_cache:
        _save(db)
    return tile, tsw, timings
