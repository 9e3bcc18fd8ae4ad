"""Autotuner for the FP32 pair kernel's launch shape.

For a workload bucket (device, atom count and q count rounded up to powers of
two) it times every candidate shape (q values per pair sweep, sine form, j
atoms per work item, persistent workgroups per CU) on the caller's structure
with the kernel's own HIP-event timing, keeps the fastest, applies it to the
context and caches it in a JSON file so later processes reuse it.  Every
candidate computes the same sums to the same accuracy; only speed differs.
"""
from __future__ import annotations

import itertools
import json
import os
from pathlib import Path
from typing import Dict, Iterable, Optional, Sequence

import numpy as np

from .api import Context, SolventParams, Structure, default_context, solvent_corrected_intensity
from . import api as _api

CACHE = Path(os.environ.get("CRABSAXS_TUNING_CACHE",
                            Path.home() / ".cache" / "crabsaxs_amd" / "tuning.json"))

DEFAULT_CANDIDATES = {
    "q_tile": (0, 64, 128),
    "sin_mode": (3,),
    "j_chunk": (256, 512, 1024),
    "persist": (0, 4, 8),
}


def _bucket(x: int) -> int:
    return 1 << max(int(x) - 1, 0).bit_length()


def workload_key(ctx: Context, natoms: int, nq: int) -> str:
    return f"{ctx.arch()}:cu{ctx.cu_count()}|atoms<={_bucket(natoms)}|q<={_bucket(nq)}"


def _load() -> Dict[str, dict]:
    try:
        return json.loads(CACHE.read_text())
    except (OSError, ValueError):
        return {}


def _store(table: Dict[str, dict]) -> None:
    try:
        CACHE.parent.mkdir(parents=True, exist_ok=True)
        tmp = CACHE.with_suffix(".tmp")
        tmp.write_text(json.dumps(table, indent=1, sort_keys=True))
        tmp.replace(CACHE)
    except OSError:
        pass


def apply(ctx: Context, shape: dict) -> None:
    ctx.set_tuning(shape["q_tile"], shape["sin_mode"], shape["j_chunk"])
    ctx.set_persistence(shape["persist"])


def _time_shapes(run, ctx: Context, cand: dict, repeats: int):
    # every timed call runs the kernel: no cached tables
    with ctx.caching(0):
        return _time_shapes_uncached(run, ctx, cand, repeats)


def _time_shapes_uncached(run, ctx: Context, cand: dict, repeats: int):
    results = []
    for qt, sm, jc, pe in itertools.product(cand["q_tile"], cand["sin_mode"], cand["j_chunk"],
                                            cand["persist"]):
        shape = {"q_tile": qt, "sin_mode": sm, "j_chunk": jc, "persist": pe}
        apply(ctx, shape)
        run()  # point layouts for this j_chunk, warm-up
        ctx.reset_timing()
        ctx.set_timing(True)
        for _ in range(repeats):
            run()
        ctx.set_timing(False)
        n, ms = ctx.kernel_time("debye_pairs_f32")
        results.append((ms / max(n, 1), shape))
    return results


def autotune_batch(structures: Sequence[Structure], q, ctx: Context = None,
                   candidates: Optional[dict] = None, sample: int = 64, repeats: int = 2,
                   params: SolventParams = None) -> dict:
    """Pick and apply the fastest pair-kernel shape for a conformer batch.

    Times csx_batch_debye_solvent_basis on an evenly spaced sample of the
    batch (sizes spread like the whole set); the chosen shape is applied to
    the context and returned with its kernel time on the sample."""
    ctx = ctx or default_context()
    q = np.asarray(q, dtype=np.float64)
    params = params or SolventParams()
    order = np.argsort([len(s) for s in structures])
    pick = [structures[int(i)] for i in order[np.linspace(0, len(order) - 1,
                                                          min(sample, len(order))).astype(int)]]
    cand = dict(DEFAULT_CANDIDATES, **(candidates or {}))
    results = _time_shapes(lambda: _api.batch_debye_solvent_basis(pick, q, params, ctx=ctx), ctx,
                           cand, repeats)
    best_ms, best = min(results, key=lambda t: t[0])
    apply(ctx, best)
    return dict(best, kernel_ms=best_ms, sample=len(pick), tried=len(results))


def autotune(structure: Structure, q, ctx: Context = None,
             candidates: Optional[dict] = None, repeats: int = 3,
             use_cache: bool = True, params: SolventParams = None) -> dict:
    """Pick and apply the fastest launch shape for this workload bucket."""
    ctx = ctx or default_context()
    q = np.asarray(q, dtype=np.float64)
    key = workload_key(ctx, len(structure), q.size)
    table = _load() if use_cache else {}
    if use_cache and key in table:
        apply(ctx, table[key])
        return dict(table[key], cached=True)
    cand = dict(DEFAULT_CANDIDATES, **(candidates or {}))
    params = params or SolventParams()
    results = _time_shapes(lambda: solvent_corrected_intensity(structure, q, params, ctx=ctx),
                           ctx, cand, repeats)
    best_ms, best = min(results, key=lambda t: t[0])
    best = dict(best, kernel_ms=best_ms)
    apply(ctx, best)
    if use_cache:
        table[key] = best
        _store(table)
    return dict(best, cached=False, tried=len(results))
