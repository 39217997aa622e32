"""Autotuners: the SELL-C spmv of the RK stage kernels and the tile shape of
the matrix-form MFMA kernels.

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


# ---- matrix-form MFMA kernels -------------------------------------------------
# The shape (tiled, nsplit, hsplit) of qa_rk_set_dense_shape: LDS-staged
# 32 x 32 workgroup tiles or one 16 x 16 tile per workgroup, and the waves
# that split each tile's K range.  Measured like the lanes: every candidate
# on the engine itself (before its state is set), interleaved rounds of the
# stage launches in solve order, the fastest round counts.
MIN_DENSE_TUNE_N = 128


def dense_candidates(N, fp32_mfma, herm, default):
    """Valid shapes for an N x N state: the default first."""
    kstep = 16 if fp32_mfma else 32
    tiled0, ns0, hs0 = default
    out = [tuple(default)]
    if N % 64 == 0:
        out.append((1, ns0, hs0))
    for s in (1, 2, 4, 8):
        if herm and N % (32 * s) == 0:
            out.append((0, ns0, s))
        if not herm and N % (kstep * s) == 0:
            out.append((0, s, hs0))
    seen, uniq = set(), []
    for c in out:
        if c not in seen:
            seen.add(c)
            uniq.append(c)
    return uniq


def dense_shape_for(eng, reps=4, rounds=3):
    """The measured best shape for the dense engine ``eng`` in its current
    Hermitian mode (None below MIN_DENSE_TUNE_N or with QUTIP_AMD_NO_TUNE);
    cached per (N, collapse operators, precision, mode)."""
    import ctypes as C

    from . import _lib
    N = int(round(eng.n ** 0.5))
    if N < MIN_DENSE_TUNE_N or os.environ.get("QUTIP_AMD_NO_TUNE"):
        return None
    herm = eng.hermitian
    nc = len(eng._dense_host[1]) // (N * N) if eng._dense_host[1].size > 1 else 0
    k = f"dense-n{N}-c{nc}-{eng.dense_precision}-{'herm' if herm else 'full'}"
    if not _CACHE:
        _load_file()
    if k not in _CACHE:
        default = eng.dense_shape
        cands = dense_candidates(N, eng.dense_precision == "fp32", herm, default)
        best = {c: float("inf") for c in cands}
        t_step, t_dense = C.c_double(), C.c_double()
        for rnd in range(rounds + 1):
            for c in cands:
                eng.dense_shape = c
                _lib.call("qa_rk_profile_sequence", eng.handle, reps, C.byref(t_step),
                          C.byref(t_dense))
                if rnd > 0:
                    best[c] = min(best[c], t_step.value)
        pick = min(best, key=best.get)
        _CACHE[k] = {"shape": list(pick),
                     "ms": {",".join(map(str, c)): v for c, v in best.items()}}
        _save_file()
    return tuple(_CACHE[k]["shape"])
