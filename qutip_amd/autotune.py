"""Autotuner for the SELL-C spmv of the RK stage kernels.

The tile shape is the number of lanes per row L (1, 2, 4, 8; C = 64 / L
rows per wavefront).  Few lanes per row keep the padding low, more lanes per
row give more wavefronts in flight (latency hiding for the column gathers).
The choice is measured: for each candidate a throw-away engine times its
stage kernels with HIP events (qa_rk_profile_sequence), and the fastest
average wins.  Results are cached per (dimension bucket, non-zeros per row
bucket, value precision, batch); the cache lives in memory and, when
QUTIP_AMD_TUNE_CACHE names a file, in that JSON file across processes.
"""
import json
import math
import os

import numpy as np

_CACHE = {}
CANDIDATES = (1, 2, 4, 8)
MIN_TUNE_ROWS = 4096  # below this the stage kernels are launch-bound


def _key(n, nnz_row, fp32, batch=1, coded=False):
    k = f"n{int(math.log2(max(n, 1)))}-r{int(round(nnz_row))}-{'c64' if fp32 else 'c128'}-b{batch}"
    return k + "-coded" if coded else k


def _load_file():
    path = os.environ.get("QUTIP_AMD_TUNE_CACHE")
    if path and os.path.exists(path):
        try:
            with open(path) as f:
                _CACHE.update(json.load(f))
        except (OSError, ValueError):
            pass


def _save_file():
    path = os.environ.get("QUTIP_AMD_TUNE_CACHE")
    if path:
        try:
            with open(path, "w") as f:
                json.dump(_CACHE, f, indent=1, sort_keys=True)
        except OSError:
            pass


def heuristic_lanes(nnz_row):
    """Default without a measurement: about 5 entries per lane."""
    L = 1
    while L < 8 and nnz_row / L > 6:
        L *= 2
    return L


def measure(A, fp32=False, candidates=CANDIDATES, reps=8, rounds=4, coded=False):
    """Average stage-kernel time (ms) per candidate lane count.  All candidate
    engines are built first, one untimed round warms the device, then the
    candidates are timed in interleaved rounds and the fastest round counts
    (so clock ramps and cache state do not favour whichever runs last).
    Each round times ``reps`` back-to-back passes over the step stages in
    solve order (qa_rk_profile_sequence)."""
    import ctypes as C

    from . import _lib
    from .solver.hip_rk import RKEngine
    engines = {L: RKEngine(A, "vern7", 1e-6, 1e-8, fp32=fp32, lanes=L, coded=coded)
               for L in candidates}
    best = {L: float("inf") for L in candidates}
    t_step, t_dense = C.c_double(), C.c_double()
    for rnd in range(rounds + 1):
        for L, eng in engines.items():
            _lib.call("qa_rk_profile_sequence", eng.handle, reps, C.byref(t_step),
                      C.byref(t_dense))
            if rnd > 0:
                best[L] = min(best[L], t_step.value)
    return best


def lanes_for(A, fp32=False, coded=False):
    n = A.shape[0]
    nnz_row = A.nnz / max(n, 1)
    if n < MIN_TUNE_ROWS or os.environ.get("QUTIP_AMD_NO_TUNE"):
        return heuristic_lanes(nnz_row)
    if not _CACHE:
        _load_file()
    k = _key(n, nnz_row, fp32, coded=coded)
    if k not in _CACHE:
        t = measure(A, fp32, coded=coded)
        _CACHE[k] = {"lanes": min(t, key=t.get), "ms": {str(L): v for L, v in t.items()}}
        _save_file()
    return int(_CACHE[k]["lanes"])


def cache():
    return dict(_CACHE)
