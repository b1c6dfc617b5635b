"""Autotuner host logic (no GPU): candidate spaces, choice of the fastest
measured shape, failing shapes skipped, and the per-device JSON cache."""
import time

import pytest

from desv import _capi, tune


class FakeBatch:
    """Stands in for desv.batch.Batch: run() takes longer for some shapes."""

    def __init__(self, cost):
        self.code = _capi.C64
        self.cost = cost
        self.config = {"tile_bits": 0, "fuse_bits": 2, "use_mfma": False, "use_graph": True}
        self.configured = []
        self.runs = 0

    def signature(self):
        return "fake-sig"

    def configure(self, tile_bits=0, use_graph=True, fuse_bits=0, use_mfma=False):
        if tile_bits == 12:
            raise ValueError("tile_bits exceeds the LDS of this device")
        self.config = {"tile_bits": tile_bits, "fuse_bits": fuse_bits or 2, "use_mfma": use_mfma,
                       "use_graph": use_graph}
        self.configured.append(dict(self.config))

    def run(self):
        self.runs += 1
        c = self.config
        time.sleep(self.cost(c["tile_bits"], c["fuse_bits"], c["use_mfma"]))


def test_candidate_spaces():
    c64 = tune.batch_candidates(_capi.C64)
    c128 = tune.batch_candidates(_capi.C128)
    assert any(c["use_mfma"] for c in c64) and not any(c["use_mfma"] for c in c128)
    assert all(1 <= c["fuse_bits"] <= 5 for c in c64 + c128)
    p = tune.program_candidates("complex64", 30)
    assert {(c["fuse_bits"], c["use_mfma"]) for c in p} >= {(2, False), (4, True), (5, True)}
    assert all(c["tile_bits"] <= 10 for c in tune.program_candidates("complex128", 10))


def test_batch_tuner_picks_fastest_and_caches(tmp_path, monkeypatch):
    monkeypatch.setenv("DESV_TUNE_CACHE", str(tmp_path / "tune.json"))
    cost = lambda t, f, m: 0.001 if (t, f, m) == (10, 4, True) else 0.004
    b = FakeBatch(cost)
    res = tune.autotune_batch(b, reps=1)
    assert res["best"] == {"tile_bits": 10, "fuse_bits": 4, "use_mfma": True}
    assert not res["cached"]
    assert b.config["tile_bits"] == 10 and b.config["use_mfma"]
    failed = [c for c in res["candidates"] if c["ms"] is None]
    assert failed and all("LDS" in c["error"] for c in failed)
    # second call: served from the cache, no timing runs
    b2 = FakeBatch(cost)
    res2 = tune.autotune_batch(b2, reps=1)
    assert res2["cached"] and b2.runs == 0 and b2.config["fuse_bits"] == 4
    # cache off: measured again
    b3 = FakeBatch(lambda t, f, m: 0.001 if t == 9 else 0.003)
    res3 = tune.autotune_batch(b3, reps=1, cache=False)
    assert res3["best"]["tile_bits"] == 9 and b3.runs > 0


def test_no_valid_candidate_raises(tmp_path, monkeypatch):
    monkeypatch.setenv("DESV_TUNE_CACHE", str(tmp_path / "tune.json"))
    with pytest.raises(RuntimeError):
        tune.autotune_batch(FakeBatch(lambda *a: 0.0), candidates=[{"tile_bits": 12, "fuse_bits": 2,
                                                                    "use_mfma": False}], cache=False)
