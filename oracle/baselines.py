"""CPU baselines of the benchmark configurations on the host's cores -- TEST
INFRASTRUCTURE (the cpu_baseline leg of bench.py), never the product.

Every worker is a separate process (spawned, so no GPU state crosses into
it) running the C restatement of the reference's CPU path single-threaded,
as the reference runs one solve per process (BLAS threads pinned to 1):
* C1 / C2: whole mesolve solves (qto_rk_solve on the Liouvillian);
* C3: mcsolve trajectories (qto_mc_traj) farmed over the workers, as
  map='parallel' farms them (qutip/solver/parallel.py:267-337);
* C4: the banded matrix form (qto_rk_solve_dia), the reference's Dia
  products, on a bounded window of the solve;
* C5: the first requests of the serving trace farmed over the workers,
  each solved on the first output interval of its time list: sesolve by
  the C RK on -iH, mesolve by the restated RK on the matrix-form RHS with
  numpy's (BLAS) dense products -- the reference's path for dense collapse
  operators (lindblad_matrix_form.pyx:170-186 over zgemm) -- and mcsolve by
  the C trajectory loop.
"""
import multiprocessing as mp
import os
import time

os.environ.setdefault("OMP_NUM_THREADS", "1")
os.environ.setdefault("OPENBLAS_NUM_THREADS", "1")
os.environ.setdefault("MKL_NUM_THREADS", "1")

import numpy as np  # noqa: E402

MAX_WORKERS = 16


def workers():
    """The cores this process may use, at most MAX_WORKERS."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, MAX_WORKERS))


def _pool(nproc):
    return mp.get_context("spawn").Pool(nproc)


# ---- workers (module level: picklable) ---------------------------------------
def _c1_worker(seconds):
    from oracle import cbind, models, tableaus
    m = models.jaynes_cummings()
    Lm = models.liouvillian(m["H"], m["c_ops"])
    fun = cbind.functionals_from_ops(m["e_ops"], 20)
    rho0 = models.vec(models.ket2dm(m["psi0"]))
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        cbind.rk_solve(Lm, rho0, m["tlist"], tableaus.vern7, fun, rtol=1e-8, atol=1e-10)
        n += 1
    return n, time.perf_counter() - t0


def _c2_worker(_):
    from oracle import cbind, models, tableaus
    m = models.ising(8, 0.1)
    Lm = models.liouvillian(m["H"], m["c_ops"])
    fun = cbind.functionals_from_ops(m["e_ops"], 256)
    rho0 = models.vec(models.ket2dm(m["psi0"]))
    t0 = time.perf_counter()
    cbind.rk_solve(Lm, rho0, m["tlist"], tableaus.vern7, fun, rtol=1e-8, atol=1e-10)
    return time.perf_counter() - t0


def _c3_worker(args):
    seeds, seconds = args
    from oracle import models, tableaus
    from oracle.cbind import McTrajectories
    m = models.ising(10, 0.05, tf=20.0, npts=100)
    cm = McTrajectories(m["H"], m["c_ops"], m["e_ops"], tableaus.TABLEAUS["vern7"])
    n, t0 = 0, time.perf_counter()
    for s in seeds:
        draws = np.random.default_rng(np.random.SeedSequence(s)).random(64)
        cm.run(m["psi0"], m["tlist"], draws)
        n += 1
        if time.perf_counter() - t0 > seconds:
            break
    return n, time.perf_counter() - t0


def _c4_worker(tf):
    from oracle import cbind, models, tableaus
    m = models.kerr(256)
    keep = m["tlist"] <= tf
    rho0 = models.vec(models.ket2dm(m["psi0"]))
    fun = cbind.functionals_from_ops(m["e_ops"], 256)
    t0 = time.perf_counter()
    _, st = cbind.rk_solve_dia(m["H"], m["c_ops"], rho0, m["tlist"][keep],
                               tableaus.TABLEAUS["vern7"], fun)
    return time.perf_counter() - t0, st["n_rhs"]


def _c5_worker(req):
    """One request of the serving mix on [tlist[0], tlist[1]]: returns the
    seconds it took."""
    from oracle import cbind, models, rk, tableaus
    from threadpoolctl import threadpool_limits
    kind, H, cs, e_ops, psi0, tl, ntraj, seed, rtol, atol = req
    tab = tableaus.TABLEAUS["vern7"]
    d = H.shape[0]
    # one BLAS thread per worker process, whatever the environment says
    threadpool_limits(1)
    t0 = time.perf_counter()
    if kind == "se":
        fun = (np.array([0, d], np.int32), np.arange(d, dtype=np.int32),
               np.ones(d, np.complex128).view(np.float64))
        cbind.rk_solve(-1j * H, psi0, tl, tab, fun, rtol=rtol, atol=atol, nsteps=10 ** 6)
    elif kind == "me":
        f = models.matrix_form_rhs(H, cs)
        rk.solve(f, models.vec(models.ket2dm(psi0)), tl, tab, rtol=rtol, atol=atol,
                 nsteps=10 ** 6)
    else:
        cm = cbind.McTrajectories(H, cs, e_ops, tab, rtol=rtol, atol=atol)
        ss = np.random.SeedSequence(seed).spawn(ntraj)
        for s_ in ss:
            cm.run(psi0, tl, np.random.default_rng(s_).random(64))
    return time.perf_counter() - t0


# ---- baselines -----------------------------------------------------------------
def c1(seconds=3.0):
    """JC mesolve solves/s: one core, and every worker solving at once."""
    n1, t1 = _c1_worker(seconds)
    nproc = workers()
    with _pool(nproc) as p:
        res = p.map(_c1_worker, [seconds] * nproc)
    agg = sum(n / t for n, t in res)
    return {"one_core": round(n1 / t1, 2), "all_cores": round(agg, 1), "cores": nproc,
            "sample": f"{seconds:.0f} s of full solves per process"}


def c2(one_core_s=None):
    """TFIM L=8 mesolve solves/s: every worker one full solve at once."""
    nproc = workers()
    t0 = time.perf_counter()
    with _pool(nproc) as p:
        ts = p.map(_c2_worker, range(nproc))
    wall = time.perf_counter() - t0
    return {"all_cores": round(nproc / max(ts), 4), "cores": nproc,
            "per_process_s": round(float(np.median(ts)), 2), "wall_s": round(wall, 2),
            "sample": f"{nproc} concurrent full solves, one per process"}


def c3(seconds=8.0, ntraj=2048):
    """TFIM L=10 mcsolve trajectories/s: one core, and the seeds 0..2047
    farmed over the workers (each stops after ``seconds``)."""
    n1, t1 = _c3_worker((list(range(ntraj)), seconds / 2))
    nproc = workers()
    chunks = [list(range(i, ntraj, nproc)) for i in range(nproc)]
    with _pool(nproc) as p:
        res = p.map(_c3_worker, [(c, seconds) for c in chunks])
    agg = sum(n / t for n, t in res)
    return {"one_core": round(n1 / t1, 2), "all_cores": round(agg, 1), "cores": nproc,
            "sample": f"the first trajectories of seeds 0..{ntraj - 1} (strided over the "
                      f"workers), {seconds:.0f} s per process; {sum(n for n, _ in res)} "
                      "trajectories"}


def c4(rhs_full, tf=0.6):
    """Kerr N=256 matrix form with the reference's banded (Dia) products:
    the window t in [0, tf] timed, the full solve's time estimated as that
    times (RHS evaluations of the full solve / of the window)."""
    t1, r1 = _c4_worker(tf)
    nproc = workers()
    with _pool(nproc) as p:
        res = p.map(_c4_worker, [tf] * nproc)
    scale = rhs_full / r1
    one = 1.0 / (t1 * scale)
    allc = sum(1.0 / (t * scale) for t, _ in res)
    return {"one_core": round(one, 5), "all_cores": round(allc, 4), "cores": nproc,
            "sample": f"t in [0, {tf}] of the window ({r1} RHS, {t1:.1f} s on one core), "
                      f"scaled by {rhs_full} RHS of the full solve"}


def c5(requests):
    """Serving-mix solves/s on the host's cores: ``requests`` (tuples of
    kind, H, c_ops, e_ops, psi0, tlist, ntraj, seed, rtol, atol) farmed over
    the workers, each solved on its first output interval; the full solves
    are estimated as that work times the number of intervals (the step
    sizes of these constant generators change little over a solve)."""
    nproc = workers()
    nint = len(requests[0][5]) - 1
    reqs = [r[:5] + (np.asarray(r[5][:2]),) + r[6:] for r in requests]
    t0 = time.perf_counter()
    with _pool(nproc) as p:
        ts = p.map(_c5_worker, reqs, chunksize=1)
    wall = time.perf_counter() - t0
    kinds = {}
    for r, t in zip(requests, ts):
        kinds.setdefault(r[0], []).append(t * nint)
    return {"all_cores": round(len(requests) / (wall * nint), 4), "cores": nproc,
            "wall_s": round(wall, 2),
            "est_full_solve_s_by_kind": {k: round(float(np.mean(v)), 2) for k, v in kinds.items()},
            "sample": f"the first {len(requests)} requests of the trace, each solved over its "
                      f"first output interval (1 of {nint}) on {nproc} workers; full solves "
                      f"estimated as {nint} x that work"}
