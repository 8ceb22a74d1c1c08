"""Autotuner for the batched assembly + PCG launch.

Candidates are timed on the actual problem (HIP events around one full
assemble + solve) and the fastest configuration is kept.  Results are cached
per (device, problem signature) in a JSON file (``$TM_TUNE_CACHE`` or
``~/.cache/tensormesh_amd/tune.json``) so later runs reuse them without
re-timing.  Tuned knobs:

* ``rows_per_tile`` - rows per PCG tile (granularity of the per-system
  reductions and of the persistent grid-stride loop),
* ``lanes`` - lanes cooperating on one CSR row in the SpMV,
* ``blocks_per_cu`` - persistent workgroups per compute unit,
* ``plan_rows`` - owned rows per assembly block (LDS tile of the assembly).
"""
from __future__ import annotations

import itertools
import json
import os
import sys
import time
from typing import Dict, Iterable, Optional

import torch

DEFAULT_SPACE = {
    "rows_per_tile": (1024, 2048, 4096),
    "lanes": (2, 4, 8),
    "blocks_per_cu": (1,),
    "plan_rows": (None,),
}


def _cache_path() -> str:
    return os.environ.get("TM_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "tensormesh_amd",
                                                        "tune.json"))


def _load() -> Dict:
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(d: Dict):
    p = _cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        with open(p + ".tmp", "w") as f:
            json.dump(d, f, indent=1)
        os.replace(p + ".tmp", p)
    except OSError:
        pass


def signature(bp) -> str:
    props = torch.cuda.get_device_properties(bp.device)
    arch = getattr(props, "gcnArchName", props.name)
    return f"{arch}|cu{props.multi_processor_count}|mesh{bp.n_mesh}|rows{bp.n_inner}|nnz{bp.nnz_inner}"


def time_config(bp, rtol: float, maxiter: int, repeats: int = 2) -> float:
    dev = bp.device
    st = torch.cuda.current_stream(dev)
    bp.run(None, bp.load_c if hasattr(bp, "load_c") else 1.0, rtol, maxiter)   # warm-up
    torch.cuda.synchronize(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    best = float("inf")
    for _ in range(repeats):
        e0.record(st)
        bp.run(None, 1.0, rtol, maxiter)
        e1.record(st)
        torch.cuda.synchronize(dev)
        if int(bp.timeout.item()) != 0:
            return float("inf")
        best = min(best, e0.elapsed_time(e1) / 1e3)
    return best


def autotune_batched(bp, space: Optional[Dict[str, Iterable]] = None, rtol: float = 1e-8, maxiter: int = 20000,
                     use_cache: bool = True, verbose: bool = False) -> Dict:
    """Pick the fastest configuration for ``bp`` (a :class:`BatchedPoisson`) and apply it."""
    space = dict(DEFAULT_SPACE, **(space or {}))
    key = signature(bp)
    cache = _load() if use_cache else {}
    if key in cache:
        cfg = cache[key]
        bp.configure(cfg["rows_per_tile"], cfg["lanes"], cfg["blocks"])
        return dict(cfg, cached=True)
    ncu = torch.cuda.get_device_properties(bp.device).multi_processor_count
    trials = []
    for rpt, lanes, bpc, prow in itertools.product(space["rows_per_tile"], space["lanes"], space["blocks_per_cu"],
                                                   space["plan_rows"]):
        if prow is not None:
            bp.plan = bp.topology.device_plan(bp.device, "scalar", prow)
        bp.configure(rpt, lanes, bpc * ncu)
        t = time_config(bp, rtol, maxiter)
        trials.append(dict(rows_per_tile=rpt, lanes=lanes, blocks=bpc * ncu, plan_rows=prow, seconds=t))
        if verbose:
            print(f"[autotune] tile {rpt} lanes {lanes} blocks/CU {bpc} plan {prow}: {t * 1e3:.2f} ms",
                  file=sys.stderr, flush=True)
    best = min(trials, key=lambda d: d["seconds"])
    if best["plan_rows"] is not None:
        bp.plan = bp.topology.device_plan(bp.device, "scalar", best["plan_rows"])
    bp.configure(best["rows_per_tile"], best["lanes"], best["blocks"])
    cfg = dict(best, trials=trials, time=time.time())
    cache[key] = cfg
    if use_cache:
        _store(cache)
    return dict(cfg, cached=False)
