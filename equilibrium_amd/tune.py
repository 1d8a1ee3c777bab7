"""Autotuner for the red-black mode: band shape x sweeps per launch.

For a scene it times one frame per configuration on the GPU and keeps the
fastest; results are cached per (size, sweeps, device name) in a JSON file
(EQF_TUNE_CACHE, default equilibrium_amd/tuning_cache.json) so later runs
reuse them.  The wavefront mode has no tile knobs: its launch shape is one
workgroup per scene and one wave per 64-row strip.
"""
import json
import os
import time

from . import _abi

# band shape index -> (waves per workgroup, rows per wave); regions are 128 columns wide
TILES = {0: (8, 16), 1: (8, 32), 2: (16, 16)}
SWEEPS_PER_BLOCK = (2, 4, 5, 6, 8)
CACHE = os.environ.get("EQF_TUNE_CACHE", os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                       "tuning_cache.json"))


def _device_name():
    try:
        import torch
        if torch.cuda.is_available():
            return torch.cuda.get_device_name(0)
    except Exception:
        pass
    return "unknown"


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
    key = f"rbband:{p.size}:{p.gs_iters}:{p.pressure_iters}:{_device_name()}"
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
