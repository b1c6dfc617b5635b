"""Autotuner for launch shapes (tile bits, fusion width, MFMA blocks).

The engine's stage kernels trade tile size (LDS per workgroup, hence
workgroups per CU and HBM passes per circuit) against fusion width (VALU /
MFMA work per amplitude and LDS round trips).  The best point depends on the
circuit, the qubit count and the batch mix, so it is measured: every
candidate shape is built and timed on the device, the fastest is kept, and
the result is cached per device and workload signature in a JSON file
(``$DESV_TUNE_CACHE``, default ``~/.cache/desv/tune.json``).

* :func:`autotune_batch` -- heterogeneous VQE batches (``desv.batch.Batch``):
  tile bits x fusion width x forward MFMA blocks, timed on full
  energy + gradient runs.
* :func:`autotune_program` -- a compiled :class:`desv.program.Program` on a
  scratch state: fusion width x tile bits x MFMA blocks.

There is no counterpart in the reference (its JAX path has no launch shapes);
the tuner only chooses among configurations that the parity tests cover.
"""
from __future__ import annotations

import json
import os
import time
from typing import Dict, Iterable, List, Optional

import numpy as np

from . import _capi

_CACHE_ENV = "DESV_TUNE_CACHE"


def _cache_path() -> str:
    return os.environ.get(_CACHE_ENV) or os.path.join(os.path.expanduser("~"), ".cache", "desv", "tune.json")


def _device_key(device: int = 0) -> str:
    try:
        import torch
        if torch.cuda.is_available():
            p = torch.cuda.get_device_properties(device)
            return f"{getattr(p, 'gcnArchName', p.name)}-{p.multi_processor_count}cu"
    except Exception:  # noqa: BLE001 - the key only scopes the cache
        pass
    return f"dev{device}"


def _load() -> Dict:
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(table: Dict) -> None:
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # read-only home: tuning still applies to this process


def batch_candidates(dtype_code: int) -> List[Dict]:
    """Default search space for batches."""
    tiles = (9, 10, 11, 12) if dtype_code == _capi.C64 else (9, 10, 11)
    out = [{"tile_bits": t, "fuse_bits": 2, "use_mfma": False} for t in tiles]
    out += [{"tile_bits": t, "fuse_bits": 3, "use_mfma": False} for t in tiles[1:3]]
    if dtype_code == _capi.C64:
        out += [{"tile_bits": t, "fuse_bits": 4, "use_mfma": True} for t in tiles[1:3]]
    return out


def program_candidates(dtype: str, n: int) -> List[Dict]:
    """Default search space for a single program."""
    c64 = np.dtype(dtype) == np.complex64
    tiles = [t for t in ((11, 12, 13) if c64 else (10, 11, 12)) if t <= n] or [n]
    out = [{"fuse_bits": f, "tile_bits": t, "use_mfma": False} for f in (2, 3) for t in tiles]
    if c64:
        out += [{"fuse_bits": f, "tile_bits": t, "use_mfma": True} for f in (3, 4, 5) for t in tiles]
        # register-resident tiles: groups (2^12 tiles), whole tiles per wave (2^11, 2^12)
        out += [{"fuse_bits": 3, "tile_bits": 12, "use_mfma": True, "reg_groups": 1}] if 12 in tiles else []
        out += [{"fuse_bits": f, "tile_bits": t, "use_mfma": True, "reg_groups": 2}
                for f in (2, 3) for t in (11, 12) if t in tiles]
    return out


def _timed_best(configs: Iterable[Dict], trial) -> Dict:
    report = []
    for cfg in configs:
        try:
            ms = trial(cfg)
        except (ValueError, MemoryError, NotImplementedError) as e:  # shape not valid here
            report.append({**cfg, "ms": None, "error": str(e)})
            continue
        report.append({**cfg, "ms": ms})
    ok = [r for r in report if r["ms"] is not None]
    if not ok:
        raise RuntimeError("autotune: no candidate configuration ran")
    best = min(ok, key=lambda r: r["ms"])
    return {"best": {k: v for k, v in best.items() if k not in ("ms", "error")}, "best_ms": best["ms"],
            "candidates": report}


def autotune_batch(batch, candidates: Optional[List[Dict]] = None, reps: int = 2, cache: bool = True) -> Dict:
    """Pick the fastest launch shape for `batch` and configure it.

    Each candidate rebuilds the batch (compile, device tables, graph capture),
    runs it once untimed and then `reps` timed energy + gradient evaluations
    (wall time of the synchronous run).  Returns {"best", "best_ms",
    "candidates", "cached"}.
    """
    key = f"batch/{_device_key()}/{batch.signature()}"
    table = _load() if cache else {}
    use_graph = batch.config.get("use_graph", True)
    if cache and key in table:
        best = table[key]["best"]
        batch.configure(use_graph=use_graph, **best)
        return {**table[key], "cached": True}

    def trial(cfg):
        batch.configure(use_graph=use_graph, **cfg)
        batch.run()
        t0 = time.perf_counter()
        for _ in range(reps):
            batch.run()
        return (time.perf_counter() - t0) / reps * 1e3

    res = _timed_best(candidates or batch_candidates(batch.code), trial)
    batch.configure(use_graph=use_graph, **res["best"])
    if cache:
        table[key] = res
        _store(table)
    return {**res, "cached": False}


def autotune_program(prog, dtype: str = "complex64", candidates: Optional[List[Dict]] = None, reps: int = 2,
                     cache: bool = True, signature: Optional[str] = None, device: int = 0,
                     compensated: bool = False) -> Dict:
    """Compile `prog` with the fastest shape among `candidates` (cached only
    when `signature` names the workload).  ``compensated`` compiles every
    candidate with compensated matrices (complex64).

    Timing runs the program on a scratch |0...0> state of the program's width
    (its own allocation; the caller's states are untouched), one untimed run
    then `reps` timed runs with per-stage events (desv_run_program_timed).
    """
    from .simulator import StateHandle
    cache = cache and signature is not None   # a result is cached only under a caller's signature
    key = f"program/{_device_key(device)}/{np.dtype(dtype).name}{'-comp' if compensated else ''}/n{prog.n}/{signature}"
    table = _load() if cache else {}
    if key in table:
        prog.compile(dtype, compensated=compensated, **table[key]["best"])
        return {**table[key], "cached": True}
    h = StateHandle(prog.n, np.dtype(dtype) == np.complex128, device)
    try:
        def trial(cfg):
            prog.compile(dtype, compensated=compensated, **cfg)
            prog.run(h)
            return float(sum(prog.run_timed(h).sum() for _ in range(reps)) / reps)

        res = _timed_best(candidates or program_candidates(dtype, prog.n), trial)
    finally:
        h.close()
    prog.compile(dtype, compensated=compensated, **res["best"])
    if cache:
        table[key] = res
        _store(table)
    return {**res, "cached": False}
