"""Launch-shape autotuner for a batch.

Two launch parameters are free: the P2G tile shape (cells per P2G block; the
sort bins particles by these tiles) and the block size of the particle
kernels (particles per block).  ``autotune`` builds a ``BatchPlan`` for every
candidate and times, with HIP events, a short checkpointed trajectory of the
batch's own particles from rest under their growth tensors (chained forward
steps, a segment recompute, chained reverse steps), keeping the fastest
median.  A candidate replaces the default shape (the engine's default tile,
256-thread blocks) only when it is at least ``MARGIN`` faster: candidates
within run-to-run noise of each other would otherwise trade places between
runs.  Results are cached per *shape class* -- dtype, model mix, advection,
log2 buckets of the particle count, the instance count and the mean grid size,
and the candidate set when it is not the full one -- in a JSON file, so a
workload is tuned once:

    tile, block = autotune(specs, pts, dtype, advect, G_list).choice

The cache file is ``$MORPHOMPM_AMD_TUNE_CACHE`` if set, else
``~/.cache/morphompm_amd/autotune.json``; ``cache=False`` skips it.
"""
import json
import math
import os
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .engine import BatchPlan, Trajectory, DEFAULT_TILE

CANDIDATE_BLOCKS = (64, 128, 256)
DEFAULT_BLOCK = 256
MARGIN = 0.02       # a candidate must beat the default shape by this fraction


def cache_path():
    p = os.environ.get("MORPHOMPM_AMD_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "morphompm_amd", "autotune.json")


def shape_class(specs, dtype, advect):
    """Cache key of a batch: what the best launch shape depends on."""
    n = sum(s.n for s in specs)
    nodes = np.mean([np.prod(s.shape) for s in specs])
    kinds = sorted({int(s.kind) for s in specs})
    dt = "f64" if dtype == torch.float64 else "f32"
    b = lambda v: int(round(math.log2(max(v, 1))))
    return (f"{dt}-m{''.join(map(str, kinds))}-a{int(bool(advect))}"
            f"-n{b(n)}-i{b(len(specs))}-g{b(nodes)}")


def cache_key(specs, dtype, advect, tiles, blocks, built):
    """The shape class, plus the candidate set when it is not every built tile
    with every candidate block (a restricted tune answers only restricted
    requests)."""
    key = shape_class(specs, dtype, advect)
    if sorted(tiles) != sorted(built) or sorted(blocks) != sorted(CANDIDATE_BLOCKS):
        key += "|" + ";".join(_label(t, b) for t in sorted(tiles) for b in sorted(blocks))
    return key


def pick(ms, default_label, margin=MARGIN):
    """The fastest candidate, unless the default is within ``margin`` of it."""
    best = min(ms, key=ms.get)
    if default_label in ms and ms[default_label] <= (1.0 + margin) * ms[best]:
        return default_label
    return best


def lookup(path, key, tiles, blocks, built):
    """The cached result for ``key`` if its shape is one of the requested
    candidates and is built, else None."""
    hit = load_cache(path).get(key)
    if not hit:
        return None
    tile = tuple(hit.get("tile", ()))
    if tile not in tiles or tile not in built or hit.get("block") not in blocks:
        return None
    return TuneResult(tile, int(hit["block"]), hit.get("ms", {}), key, True)


def load_cache(path):
    try:
        with open(path) as f:
            d = json.load(f)
        return d if isinstance(d, dict) else {}
    except (OSError, ValueError):
        return {}


def store_cache(path, key, entry):
    """Adds one entry; never raises (a read-only home just skips caching)."""
    try:
        d = load_cache(path)
        d[key] = entry
        os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
        tmp = f"{path}.{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
        return True
    except OSError:
        return False


@dataclass
class TuneResult:
    tile: tuple
    block: int
    ms: dict = field(default_factory=dict)   # "tx,ty,tz/block" -> ms per short trajectory
    key: str = ""
    cached: bool = False

    @property
    def choice(self):
        return self.tile, self.block


def _label(tile, block):
    return f"{tile[0]},{tile[1]},{tile[2]}/{block}"


def _time_candidate(specs, pts_list, G_list, dtype, advect, tile, block, steps, reps):
    """Median ms of a checkpointed trajectory of `steps` steps (chained
    forward, segment recompute, chained reverse) -- the kernels and launch mix
    a batched trajectory runs."""
    plan = BatchPlan(specs, tile=tile, block=block)
    perm = plan.cell_order(pts_list)
    parts = []
    for s, p in zip(specs, pts_list):
        k = s.n
        parts.append((p, np.zeros((k, 3)), np.zeros((k, 3, 3)), np.repeat(np.eye(3)[None], k, 0)))
    dyn0 = plan.pack_dyn(parts, dtype, perm)
    G = plan.pack_mat(G_list, dtype, perm)
    traj = Trajectory(plan, dtype, steps, advect, max(1, steps // 2))
    bar = plan.new_dyn(dtype)
    bar[15:24] = plan.live.to(dtype)           # a unit F cotangent
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ms = []
    for r in range(reps + 1):
        e0.record()
        traj.forward(dyn0, G)
        traj.backward(bar)
        e1.record()
        torch.cuda.synchronize()
        if r:                                  # the first run warms up
            ms.append(e0.elapsed_time(e1))
    return float(np.median(ms))


def autotune(specs, pts_list, dtype=torch.float32, advect=True, G_list=None, tiles=None,
             blocks=CANDIDATE_BLOCKS, steps=6, reps=5, cache=True, path=None):
    """Best (tile, block) for this batch; see the module docstring."""
    _lib.require_gpu()
    path = path or cache_path()
    built = [tuple(t) for t in _lib.tile_shapes()]
    tiles = [tuple(t) for t in (tiles or built)]
    blocks = tuple(int(b) for b in blocks)
    key = cache_key(specs, dtype, advect, tiles, blocks, built)
    if cache:
        hit = lookup(path, key, tiles, blocks, built)
        if hit is not None:
            return hit
    if G_list is None:
        G_list = [np.repeat((1.05 * np.eye(3))[None], s.n, 0) for s in specs]
    ms = {}
    t0 = time.perf_counter()
    for tile in tiles:
        for block in blocks:
            ms[_label(tile, block)] = _time_candidate(specs, pts_list, G_list, dtype, advect, tile,
                                                      block, steps, reps)
    best = pick(ms, _label(DEFAULT_TILE, DEFAULT_BLOCK))
    tl, bl = best.split("/")
    res = TuneResult(tuple(int(v) for v in tl.split(",")), int(bl), ms, key, False)
    if cache:
        store_cache(path, key, {"tile": list(res.tile), "block": res.block, "ms": ms,
                                "tuned_s": round(time.perf_counter() - t0, 3)})
    return res
