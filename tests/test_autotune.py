"""Launch-shape autotuner: shape classes and the cache on the CPU; tuning a
small batch on the GPU."""
import json
import os

import numpy as np
import pytest
import torch

from morphompm_amd import autotune as AT
from morphompm_amd.engine import InstanceSpec


def _specs(ns, shape=(32, 32, 32), kind=0):
    return [InstanceSpec(n, shape, 1 / 32, 1e-4, 0.05, 1.0, 1.0, kind, (1.0,) * 6) for n in ns]


def test_shape_class_buckets():
    a = AT.shape_class(_specs([1000, 2000]), torch.float32, True)
    # same log2 buckets -> same class; dtype, model, advect and size change it
    assert AT.shape_class(_specs([1100, 2100]), torch.float32, True) == a
    assert AT.shape_class(_specs([1000, 2000]), torch.float64, True) != a
    assert AT.shape_class(_specs([1000, 2000], kind=1), torch.float32, True) != a
    assert AT.shape_class(_specs([1000, 2000]), torch.float32, False) != a
    assert AT.shape_class(_specs([8000, 16000]), torch.float32, True) != a
    assert AT.shape_class(_specs([1000] * 8), torch.float32, True) != a


def test_cache_roundtrip(tmp_path):
    p = str(tmp_path / "sub" / "tune.json")
    assert AT.load_cache(p) == {}
    assert AT.store_cache(p, "k1", {"tile": [4, 8, 8], "block": 128, "ms": {"x": 1.0}})
    assert AT.store_cache(p, "k2", {"tile": [4, 4, 8], "block": 64, "ms": {}})
    d = AT.load_cache(p)
    assert d["k1"]["block"] == 128 and d["k2"]["tile"] == [4, 4, 8]
    with open(p, "w") as f:
        f.write("{not json")
    assert AT.load_cache(p) == {}


def test_cache_path_env(monkeypatch, tmp_path):
    monkeypatch.setenv("MORPHOMPM_AMD_TUNE_CACHE", str(tmp_path / "t.json"))
    assert AT.cache_path() == str(tmp_path / "t.json")
    monkeypatch.delenv("MORPHOMPM_AMD_TUNE_CACHE")
    assert AT.cache_path().endswith(os.path.join("morphompm_amd", "autotune.json"))


def test_cache_key_covers_candidate_set():
    built = [(4, 8, 8), (4, 4, 8), (2, 4, 8)]
    sp = _specs([1000, 2000])
    full = AT.cache_key(sp, torch.float32, True, built, AT.CANDIDATE_BLOCKS, built)
    assert full == AT.shape_class(sp, torch.float32, True)
    one = AT.cache_key(sp, torch.float32, True, [(4, 4, 8)], AT.CANDIDATE_BLOCKS, built)
    blk = AT.cache_key(sp, torch.float32, True, built, (256,), built)
    assert len({full, one, blk}) == 3


def test_lookup_accepts_only_requested_shapes(tmp_path):
    """A pre-seeded cache: a hit is returned only for a built tile the caller
    asked for, with a block the caller asked for."""
    p = str(tmp_path / "tune.json")
    built = [(4, 8, 8), (4, 4, 8), (2, 4, 8)]
    AT.store_cache(p, "k", {"tile": [4, 8, 8], "block": 128, "ms": {"4,8,8/128": 1.0}})
    r = AT.lookup(p, "k", built, AT.CANDIDATE_BLOCKS, built)
    assert r is not None and r.cached and r.choice == ((4, 8, 8), 128)
    assert AT.lookup(p, "k", [(4, 4, 8)], AT.CANDIDATE_BLOCKS, built) is None
    assert AT.lookup(p, "k", built, (256,), built) is None
    assert AT.lookup(p, "k", built, AT.CANDIDATE_BLOCKS, [(4, 4, 8)]) is None
    assert AT.lookup(p, "other", built, AT.CANDIDATE_BLOCKS, built) is None


def test_pick_keeps_default_within_margin():
    d = AT._label(AT.DEFAULT_TILE, AT.DEFAULT_BLOCK)
    assert AT.pick({d: 1.01, "2,4,8/256": 1.0}, d) == d
    assert AT.pick({d: 1.05, "2,4,8/256": 1.0}, d) == "2,4,8/256"
    assert AT.pick({"2,4,8/256": 1.0, "4,8,8/64": 1.2}, d) == "2,4,8/256"


@pytest.mark.gpu
def test_autotune_small_batch(tmp_path):
    from morphompm_amd import _lib
    from morphompm_amd.batched import BatchedInverse
    from morphompm_amd.scenes import batch_instances
    sc = batch_instances(count=4, seed=3, n_lo=500, n_hi=3000, n_steps=8)
    for s in sc:
        s.checkpoint_every = 4
    specs = [InstanceSpec.from_cfg(s.n, s.cfg, s.model, s.pts, center=True) for s in sc]
    path = str(tmp_path / "tune.json")
    r = AT.autotune(specs, [s.pts for s in sc], torch.float32, True, path=path)
    assert not r.cached
    assert r.tile in [tuple(t) for t in _lib.tile_shapes()] and r.block in AT.CANDIDATE_BLOCKS
    assert len(r.ms) == len(_lib.tile_shapes()) * len(AT.CANDIDATE_BLOCKS)
    best = min(r.ms.values())
    assert all(v > 0 for v in r.ms.values())
    assert r.ms[AT._label(r.tile, r.block)] == best or (
        (r.tile, r.block) == (AT.DEFAULT_TILE, AT.DEFAULT_BLOCK)
        and r.ms[AT._label(r.tile, r.block)] <= (1 + AT.MARGIN) * best)
    r2 = AT.autotune(specs, [s.pts for s in sc], torch.float32, True, path=path)
    assert r2.cached and r2.choice == r.choice
    assert json.load(open(path))[r.key]["block"] == r.block
    # a tuned batch computes what a fixed launch shape computes (FP64)
    os.environ["MORPHOMPM_AMD_TUNE_CACHE"] = str(tmp_path / "tune64.json")
    try:
        rates = [s.extra["rate"] for s in sc]
        ba = BatchedInverse(sc, dtype=torch.float64, tile="auto", block="auto", use_graphs=False)
        ba.set_theta(rates)
        La, ga = (t.cpu().numpy() for t in ba.loss_and_grad())
        bf = BatchedInverse(sc, dtype=torch.float64, use_graphs=False)
        bf.set_theta(rates)
        Lf, gf = (t.cpu().numpy() for t in bf.loss_and_grad())
    finally:
        del os.environ["MORPHOMPM_AMD_TUNE_CACHE"]
    assert ba.tune is not None
    np.testing.assert_allclose(La, Lf, rtol=1e-12)
    np.testing.assert_allclose(ga, gf, rtol=1e-10)
