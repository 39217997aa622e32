"""Benchmark of the time-evolution hot path on one MI355X.

Workload (one "step" = one complete solve): mesolve of the dissipative
transverse-field Ising chain, L=8 spins (d=256, Liouvillian 65536 x 65536
CSR with 1 245 183 non-zeros, complex128), J=1, h_i ~ U[0.5, 1.5]
(default_rng(0)), collapse sqrt(0.1) sigma^-_i on every site,
rho0 = |up..up><up..up|, t in [0, 10] with 200 output times, e_ops
sum(sigma_z)/L and sigma_z1 sigma_z2, vern7 at rtol=1e-8, atol=1e-10, run by
the fused hip_rk engine with inputs resident in HBM.

Printed: one JSON line with the throughput (solves/s), the roofline of the
dominant kernel (the RK stage kernel: SELL-64 spmv + stage-combine
epilogue), the CPU baseline (the C restatement of the reference CPU path,
one core, the same solve) and the <O> error against the FP64 fixture.
"""
import argparse
import ctypes as C
import json
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "mcsolve traj/s & mesolve solves/s at ≤1e-6 err vs FP64; p50/p99 latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stage_bytes(engine, nnz, n, fp32):
    """Algorithmic HBM bytes of each stage-kernel launch: the operator once
    (plain: a value and a column per non-zero; coded: a 2-byte code per
    non-zero and the two 256-entry tables), the input vector once, k_i
    written, plus the epilogue's reads of y_prev and of every k_j with a
    non-zero coefficient and its write of the next stage input (or
    y_front)."""
    tab = engine.tableau
    a, b, e = tab["a"], tab["b"], tab["e"]
    s, se = len(b), len(tab["c"])
    if engine.coded:
        op = nnz * 2 + 256 * (16 + 4)
    else:
        op = nnz * ((8 if fp32 else 16) + 4)
    base = op + (engine.sell.nslices + 1) * 4 + n * 16 + n * 16
    out = []
    for i in range(se):
        if i < s - 1 or (s <= i < se - 1):
            terms = sum(1 for j in range(i) if a[i + 1, j] != 0)
            epi = (terms + 1) * n * 16 + n * 16
        elif i == s - 1:
            terms = sum(1 for j in range(i) if b[j] != 0 or e[j] != 0)
            epi = (terms + 1) * n * 16 + n * 16
        else:
            epi = 0
        out.append(base + epi)
    return np.array(out, dtype=np.float64)


def cpu_baseline(L_csr, rho0, tlist, e_ops_funcs):
    from oracle import cbind, tableaus
    t0 = time.perf_counter()
    vals, _, st = cbind.rk_solve(L_csr, rho0, tlist, tableaus.vern7, e_ops_funcs,
                                 rtol=1e-8, atol=1e-10)
    dt = time.perf_counter() - t0
    return dt, vals, st


def mc_cpu_baseline():
    """The C restatement of the reference trajectory loop (qto_mc_traj) on
    the same workload: one core, and the seeds farmed over the host's cores
    as map='parallel' farms them."""
    from oracle import baselines
    r = baselines.c3()
    return {"value": r["all_cores"], "unit": "traj/s", "cores": r["cores"], "kind": "port",
            "one_core": r["one_core"], "sample": r["sample"]}


def bench_mcsolve(qa, systems, log, ntraj=2048, reps=3, cpu=True):
    """Batched mcsolve: TFIM L=10 (d=1024, H_eff 11264 nnz, 10 collapse
    channels), 2048 trajectories with seeds 0..2047, t in [0,20], 100 outputs,
    vern7 at the integrator defaults; complex128 and complex64 (FP64 norms).
    Timed: the device run with the pre-drawn random numbers resident in HBM
    (drawing them on the host is reported separately).  Ensemble means are
    checked against the FP64 master-equation fixture (3 sigma)."""
    from qutip_amd import _lib, data as qd
    from qutip_amd.solver.mcsolve import _read_seeds, draws_for
    s = systems.ising(10, 0.05, tf=20.0, npts=100)
    golden = np.load(ROOT / "tests" / "golden" / "ising10_mesolve.npz")
    tlist = golden["tlist"]
    out = {"workload": "mcsolve dissipative TFIM L=10 (d=1024, H_eff 11264 nnz, 10 channels), "
                       "2048 trajectories seeds 0..2047, t in [0,20], 100 outputs, vern7 "
                       "atol=1e-8 rtol=1e-6 (integrator defaults), one persistent kernel"}
    seeds = _read_seeds(list(range(ntraj)), ntraj)
    t0 = time.perf_counter()
    draws = draws_for(seeds, 64)
    out["host_draws_s"] = round(time.perf_counter() - t0, 4)
    dev_draws = (qd.DeviceBuffer.from_array(draws), ntraj, 64)
    psi0 = s["psi0"].full().ravel()
    for prec in ("fp64", "fp32"):
        solver = qa.MCSolver(s["H"], s["c_ops"], options={"precision": prec})
        solver.run_batch(psi0, tlist, s["e_ops"], dev_draws)  # warm-up
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            ex, info, _, _ = solver.run_batch(psi0, tlist, s["e_ops"], dev_draws)
            _lib.call("qa_device_sync")
            ts.append(time.perf_counter() - t0)
        best = min(ts)
        ok = bool((info[:, 0] == 0).all())
        worst_sigma = 0.0
        for k in range(2):
            v = ex[:, :, k].real
            mean, sig = v.mean(0), v.std(0) / np.sqrt(ntraj)
            dev = np.abs(mean - golden["expect"][k])
            worst_sigma = max(worst_sigma, float((dev[1:] / sig[1:]).max()))
        # operator bytes the trajectories read from L2: at least 10 RHS per
        # attempt, each reading the whole ELL (width x d columns and values)
        width = solver._device_ops(prec == "fp32")[1][0]
        vb = 8 if prec == "fp32" else 16
        op_bytes = 10 * float(info[:, 3].sum()) * width * s["H"].shape[0] * (vb + 4)
        # the algorithmic bytes of a batched RHS: every trajectory's state
        # read and its derivative written (2 d 16 B, complex128; 2 d 8 B for
        # complex64) plus the operator once per batch of ntraj; at least 10
        # RHS per attempt (the vern7 step stages; dense-output stages add more)
        rhs = 10 * float(info[:, 3].sum())
        d = s["H"].shape[0]
        alg = rhs * 2 * d * vb + rhs / ntraj * (solver.H_eff.nnz * 20 + (d + 1) * 4)
        achieved = alg / best / 1e9
        out[prec] = {"traj_per_s": round(ntraj / best, 1), "s_per_batch": round(best, 4),
                     "all_ok": ok, "jumps": int(info[:, 1].sum()),
                     "rk_attempts": int(info[:, 3].sum()),
                     "max_dev_over_sigma_vs_mesolve": round(worst_sigma, 3),
                     "ell_width": int(width),
                     "operator_l2_read_gbs_lower_bound": round(op_bytes / best / 1e9, 1),
                     "roofline": {
                         "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                         "traffic": read_mc_traffic(prec),
                         "kernel": "mc_traj_kernel (one launch, every trajectory)",
                         "avg_launch_us": round(best * 1e6, 1),
                         "algorithmic_bytes_per_launch": int(alg),
                         "algorithmic_unit": "per batched RHS: the states read and written "
                                             "(2 d 16 B per trajectory) and the operator once; "
                                             f"{int(rhs)} trajectory RHS (10 per attempt, a "
                                             "lower bound), so achieved is a lower bound; the "
                                             "launch time is the host-timed batch"}}
        log(f"mcsolve {prec}: {ntraj / best:.0f} traj/s ({best * 1e3:.1f} ms), worst "
            f"{worst_sigma:.2f} sigma, jumps {int(info[:, 1].sum())}")
    # the plain drop-in form: qutip's trajectory loop through the hip_rk
    # integrator and device data ops (MCSolver.run_plain), on the first
    # trajectories of the same draws
    solver = qa.MCSolver(s["H"], s["c_ops"], options={"precision": "fp64"})
    nplain = 4
    t0 = time.perf_counter()
    runs, jumps = solver.run_plain(psi0, tlist, s["e_ops"], draws[:nplain])
    tpl = time.perf_counter() - t0
    ex, info, _, _ = solver.run_batch(psi0, tlist, s["e_ops"], draws[:nplain])
    out["plain_form"] = {"traj_per_s": round(nplain / tpl, 3), "trajectories": nplain,
                         "same_jumps_as_batched": bool((jumps == info[:, 1]).all()),
                         "max_abs_diff_vs_batched": float(np.abs(runs - ex).max()),
                         "batched_fp64_over_plain": round(out["fp64"]["traj_per_s"] * tpl / nplain,
                                                          1)}
    log(f"mcsolve plain form: {nplain / tpl:.2f} traj/s, same jumps {out['plain_form']['same_jumps_as_batched']}, "
        f"max diff {out['plain_form']['max_abs_diff_vs_batched']:.1e}")
    if cpu:
        out["cpu_baseline"] = mc_cpu_baseline()
        log(f"mcsolve CPU baseline: {out['cpu_baseline']}")
    return out


def bench_jc(qa, systems, log, reps=20, cpu=True):
    """The CPU-runnable plain mesolve: Jaynes-Cummings N=10 x qubit (d=20,
    Liouvillian 400 x 400, 1543 non-zeros), t in [0,25] with 100 outputs,
    vern7 rtol=1e-8 atol=1e-10; fused hip_rk (a launch-bound size) beside
    the C restatement on one core."""
    from qutip_amd import _lib
    s = systems.jaynes_cummings()
    g = np.load(ROOT / "tests" / "golden" / "jc_mesolve.npz")
    tl = g["tlist"]  # read from the archive once, outside the timed loop
    solver = qa.MESolver(s["H"], s["c_ops"], options={"rtol": 1e-8, "atol": 1e-10})
    res = solver.run(s["psi0"], tl, e_ops=s["e_ops"])
    _lib.call("qa_device_sync")
    t0 = time.perf_counter()
    for _ in range(reps):
        res = solver.run(s["psi0"], tl, e_ops=s["e_ops"])
    _lib.call("qa_device_sync")
    dt = (time.perf_counter() - t0) / reps
    err = max(np.abs(np.asarray(res.expect[k]) - g["expect"][k]).max() for k in range(2))
    out = {"workload": "mesolve Jaynes-Cummings N=10 x qubit (Liouvillian 400^2, 1543 nnz), "
                       "t in [0,25], 100 outputs, vern7 rtol=1e-8 atol=1e-10",
           "solves_per_s": round(1 / dt, 2), "err_vs_fp64": float(err),
           "rk_attempts": solver._integrator.stats()["attempts"],
           "path": solver._integrator.stats().get("path", "fused engine")}
    # a sweep: 2048 couplings g in [0.02, 0.09] x 2 pi, one launch, one
    # workgroup per solve
    from qutip_amd.solver.small import SmallSolve
    gs = 2 * np.pi * np.linspace(0.02, 0.09, 2048)
    Ls = []
    for gg in gs:
        sg = systems.jaynes_cummings(g=gg)
        Ls.append(qa.liouvillian(sg["H"], sg["c_ops"]).to_csr())
    ss = SmallSolve(Ls, "vern7", rtol=1e-8, atol=1e-10)
    ss.set_functionals(*solver._functionals(s["e_ops"]))
    rho0 = np.tile(qa.ket2dm(s["psi0"]).full().ravel(order="F"), (len(gs), 1))
    ss.run(0.0, rho0[:256], tl[1:])  # warm-up
    t0 = time.perf_counter()
    vals, _, info = ss.run(0.0, rho0, tl[1:])
    dts = time.perf_counter() - t0
    out["sweep_2048_couplings"] = {
        "solves_per_s": round(len(gs) / dts, 1), "s": round(dts, 4),
        "all_ok": bool((info[:, 0] >= 0).all()),
        "rk_attempts_total": int(info[:, 3].sum())}
    if cpu:
        from oracle import baselines
        r = baselines.c1()
        out["cpu_baseline"] = {"value": r["all_cores"], "unit": "solves/s", "cores": r["cores"],
                               "kind": "port", "one_core": r["one_core"],
                               "sample": r["sample"]}
    log(f"jc: {out}")
    return out


def bench_dense(qa, systems, log, reps=2, tf=None, cpu=True):
    """Matrix-form mesolve of the driven-dissipative Kerr oscillator, N=256
    (dense rho 256 x 256), t in [0, 30] with 150 outputs, e_ops a^dag a and
    parity, vern7 at rtol=1e-8 atol=1e-10, on MFMA in FP32 and in the bf16 x6
    and x3 splits (FP32 accumulate), and in FP64 from the diagonals; error
    against the FP64 fixture; the CPU baseline is the reference's banded
    (Dia) form restated in C on a window of the solve.  Algorithmic FLOP
    per RHS: the Hermitian form's three complex 256^3 products H_nh X, c X
    and (c X) c^dag (3 * 8 * 256^3 = 402.7 MFLOP), which the reference's
    A + A^dag computes for a Hermitian rho."""
    from qutip_amd import _lib
    s = systems.kerr(256)
    g = np.load(ROOT / "tests" / "golden" / "kerr_full_mesolve.npz")
    tl = g["tlist"]
    ref = g["expect"]
    if tf is not None:  # a shorter window (profiling runs)
        keep = tl <= tf
        tl, ref = tl[keep], ref[:, keep]
    N = 256
    flop_rhs = 3 * 8 * N ** 3
    out = {"workload": "mesolve Kerr oscillator N=256 matrix form (dense rho 256x256), "
                       f"t in [0,{tl[-1]:g}], {len(tl)} outputs, vern7 rtol=1e-8 atol=1e-10",
           "algorithmic_flop_per_rhs": flop_rhs}
    for prec in ("fp32", "bf16x6", "bf16x3", "fp64_banded"):
        solver = qa.MESolver(s["H"], s["c_ops"], options={
            "matrix_form": True, "dense_precision": prec, "rtol": 1e-8, "atol": 1e-10})
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            res = solver.run(s["psi0"], tl, e_ops=s["e_ops"])
            _lib.call("qa_device_sync")
            ts.append(time.perf_counter() - t0)
        err = max(np.abs(np.asarray(res.expect[k]) - ref[k]).max() for k in range(2))
        st = solver._integrator.stats()
        eng = solver._integrator._engine
        t_step, t_dense = C.c_double(), C.c_double()
        _lib.call("qa_rk_profile_sequence", eng.handle, 10, C.byref(t_step), C.byref(t_dense))
        rhs_ms = t_step.value
        tflops = flop_rhs / (rhs_ms * 1e-3) / 1e12
        best = min(ts)
        out[prec] = {"solves_per_s": round(1 / best, 4), "s_per_solve": round(best, 3),
                     "err_vs_fp64": float(err), "rk_attempts": st["attempts"],
                     "rhs_evals": st["rhs_evals"], "stage_us": round(rhs_ms * 1e3, 2)}
        if prec != "fp64_banded":
            out[prec]["hermitian_tiles"] = eng.hermitian
        out[prec]["meets_1e-6"] = bool(err <= 1e-6)
        if prec == "fp64_banded":
            # Dia products of the diagonals (the reference's own form for this
            # system): no MFMA; the state is read once per stage
            out[prec]["path"] = "FP64 products of the operators' diagonals"
            log(f"dense {prec}: {1 / best:.3f} solves/s err {err:.2e} stage {rhs_ms * 1e3:.1f} us")
            continue
        out[prec]["mfma_tflops"] = round(tflops, 2)
        out[prec]["mfma_peak_tflops"] = 157.3 if prec == "fp32" else 2500.0
        if prec.startswith("bf16"):
            out[prec]["hardware_products_per_real_product"] = 3 if prec == "bf16x3" else 6
        log(f"dense {prec}: {1 / best:.3f} solves/s err {err:.2e} stage {rhs_ms * 1e3:.1f} us "
            f"({tflops:.1f} TFLOP/s)")
    # the same solve through the sparse Liouvillian (the banded structure
    # the reference's Dia products exploit): 65536^2 superoperator, SELL-C
    solver = qa.MESolver(s["H"], s["c_ops"], options={"rtol": 1e-8, "atol": 1e-10})
    t0 = time.perf_counter()
    res = solver.run(s["psi0"], tl, e_ops=s["e_ops"])
    _lib.call("qa_device_sync")
    dt = time.perf_counter() - t0
    err = max(np.abs(np.asarray(res.expect[k]) - ref[k]).max() for k in range(2))
    Lk = solver.rhs.qobj.to_csr()
    out["superoperator_sparse_fp64"] = {
        "solves_per_s": round(1 / dt, 4), "err_vs_fp64": float(err), "nnz": int(Lk.nnz),
        "rk_attempts": solver._integrator.stats()["attempts"]}
    log(f"kerr sparse superoperator: {1 / dt:.3f} solves/s err {err:.2e} (nnz {Lk.nnz})")
    if cpu and tf is None:
        from oracle import baselines
        r = baselines.c4(out["fp64_banded"]["rhs_evals"], tf=2.0)
        out["cpu_baseline"] = {"value": r["all_cores"], "unit": "solves/s", "cores": r["cores"],
                               "kind": "port", "one_core": r["one_core"],
                               "path": "the reference's banded (Dia) matrix-form products",
                               "sample": r["sample"]}
        log(f"kerr CPU baseline: {out['cpu_baseline']}")
    return out


def served_errors(done, log):
    """The served <O>(t) of the first requests against FP64 references, as
    tests/test_gpu_serve.py checks them: sesolve against the matrix
    exponential (scipy), mesolve against the committed FP64 fixture,
    mcsolve as |mean - FP64| / (4 se + 1e-6), at most 1 within the bar,
    the standard error bounded below by 2 ||O|| sqrt(p (1 - p) / ntraj) with
    p the no-jump probability (most requests of this mix see no jump, so
    their sample standard error is 0)."""
    import scipy.sparse.linalg as spla
    g = np.load(ROOT / "tests" / "golden" / "serving_seed7.npz")
    nref = int(g["nreq"])
    ref = {int(i): e for i, e in zip(g["ids"], g["expect"])}
    worst = {"se": 0.0, "me": 0.0, "mc_dev_over_4se_plus_1e-6": 0.0}
    for r in done:
        if r.id >= nref or r.error:
            continue
        got = np.asarray(r.result).real
        if r.kind == "se":
            psis = spla.expm_multiply(-1j * r.H.to_csr().tocsc(),
                                      r.psi0.full().ravel().astype(complex),
                                      start=r.tlist[0], stop=r.tlist[-1], num=len(r.tlist),
                                      endpoint=True)
            want = np.array([[np.vdot(p, O.to_csr() @ p).real for p in psis] for O in r.e_ops])
            worst["se"] = max(worst["se"], float(np.abs(got - want).max()))
        elif r.kind == "me":
            worst["me"] = max(worst["me"], float(np.abs(got - ref[r.id]).max()))
        else:
            H = r.H.full()
            heff = -1j * H - 0.5 * sum(c.full().conj().T @ c.full() for c in r.c_ops)
            psis = spla.expm_multiply(heff, r.psi0.full().ravel().astype(complex), start=0.0,
                                      stop=r.tlist[-1], num=len(r.tlist), endpoint=True)
            p = np.clip(1 - np.sum(np.abs(psis) ** 2, axis=1), 0.0, 1.0)
            onorm = np.array([1.0, np.linalg.norm(H, 2)])[:, None]
            se = np.maximum(np.asarray(r.stderr), 2 * onorm * np.sqrt(p * (1 - p) / r.ntraj))
            dev = np.abs(got - ref[r.id]) / (4 * se + 1e-6)
            worst["mc_dev_over_4se_plus_1e-6"] = max(worst["mc_dev_over_4se_plus_1e-6"],
                                                     float(dev.max()))
    log(f"served error (first {nref} requests): {worst}")
    return {"requests_checked": nref, **{k: float(f"{v:.3e}") for k, v in worst.items()}}


def bench_serving(systems, log, nreq=2040, warm=40, policy="sjf", cpu=True, cpu_sample=40):
    """The serving mix: Poisson arrivals at 200 req/s (seed 7), d uniform in
    {16..512}, 40% sesolve / 40% mesolve / 20% mcsolve at rtol 1e-8 /
    atol 1e-10 (the tolerances of the <O> error bar), admitted by HBM
    budget and slots, overlapped on per-request streams.  Latency =
    completion - arrival; the first ``warm`` requests are excluded from the
    percentiles.  The served <O> of the first requests is checked against
    FP64 references."""
    from qutip_amd.serve import SolveServer
    t0 = time.perf_counter()
    reqs = systems.serving_trace(nreq)
    tbuild = time.perf_counter() - t0
    # warm the kernels and the autotuner cache on a short prefix
    SolveServer(max_slots=64).run(systems.serving_trace(12), realtime=False)
    srv = SolveServer(max_slots=64, policy=policy)
    t0 = time.perf_counter()
    done = srv.run(reqs, realtime=True)
    wall = time.perf_counter() - t0
    done = sorted(done, key=lambda r: r.id)
    lat = np.array([r.latency() for r in done[warm:]])
    service = np.array([r.t_done - r.t_admit for r in done[warm:]])
    errors = [r.id for r in done if r.error]
    kinds = {k: sum(1 for r in done if r.kind == k) for k in ("se", "me", "mc")}
    by_kind = {}
    for k in ("se", "me", "mc"):
        sv = [r.t_done - r.t_admit for r in done[warm:] if r.kind == k]
        if sv:
            by_kind[k] = {"p50_service_s": round(float(np.percentile(sv, 50)), 4),
                          "p99_service_s": round(float(np.percentile(sv, 99)), 4)}
    out = {"workload": f"{nreq} requests ({warm} of warm-up), Poisson 200 req/s seed 7, "
                       "d in {16..512}, 40/40/20 se/me/mc, t_final 10, 50 outputs, "
                       "rtol 1e-8 atol 1e-10",
           "admission": policy,
           "solves_per_s": round(len(done) / wall, 2), "wall_s": round(wall, 3),
           "arrival_span_s": round(reqs[-1].arrival, 3),
           "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
           "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
           "p50_service_s": round(float(np.percentile(service, 50)), 4),
           "p99_service_s": round(float(np.percentile(service, 99)), 4),
           "service_by_kind": by_kind,
           "peak_in_flight": srv.peak_in_flight, "errors": len(errors), "kinds": kinds,
           "trace_build_s": round(tbuild, 2),
           "served_error": served_errors(done, log)}
    log(f"serving: {out['solves_per_s']} solves/s p50 {out['p50_latency_s']} s "
        f"p99 {out['p99_latency_s']} s peak {srv.peak_in_flight}")
    if cpu:
        # the reference's CPU paths (restated) for the first requests of the
        # trace on the host's cores
        from oracle import baselines
        sample = [(r.kind, r.H.to_csr(), [c.full() for c in r.c_ops],
                   [O.to_csr() for O in r.e_ops], r.psi0.full().ravel(), r.tlist, r.ntraj,
                   r.seed, r.options["rtol"], r.options["atol"]) for r in reqs[:cpu_sample]]
        b = baselines.c5(sample)
        out["cpu_baseline"] = {
            "value": b["all_cores"], "unit": "solves/s", "cores": b["cores"], "kind": "port",
            "path": "sesolve: C RK on -iH; mesolve: RK on the matrix form with numpy's dense "
                    "products (zgemm, one BLAS thread per process); mcsolve: C trajectory loop",
            "est_full_solve_s_by_kind": b["est_full_solve_s_by_kind"], "sample": b["sample"]}
        log(f"serving CPU baseline: {out['cpu_baseline']}")
    return out


COPY_BYTES = 64 << 20
TRAFFIC_FILE = ROOT / "profiles" / "rk_stage_kernel_ising8_traffic.txt"


def pmc_workload():
    """The launches the HBM-traffic counters are read on (run under
    rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE): three device-to-device
    copies of 64 MiB (the calibration: a plain copy reads and writes exactly
    that many bytes), then, for the plain and then the coded operator form
    and each SELL shape (1, 2, 4, 8 lanes per row), 3 launches of each of
    the 16 stage kernels of the L=8 Ising solve in solve order (direct
    launches, no graph)."""
    from qutip_amd import _lib, data as qd
    from qutip_amd.solver.hip_rk import RKEngine
    from tests import systems
    import qutip_amd as qa
    s = systems.ising(8, 0.1)
    L = qa.MESolver(s["H"], s["c_ops"]).rhs.qobj.to_csr()
    a, b = qd.DeviceBuffer(COPY_BYTES), qd.DeviceBuffer(COPY_BYTES)
    _lib.call("qa_fill_zero", C.c_void_p(a.ptr), C.c_size_t(COPY_BYTES), None)
    for _ in range(3):
        _lib.call("qa_memcpy_d2d", C.c_void_p(b.ptr), C.c_void_p(a.ptr), C.c_size_t(COPY_BYTES),
                  None)
    _lib.call("qa_device_sync")
    rho0 = np.zeros(L.shape[0], dtype=np.complex128)
    rho0[0] = 1.0
    ms = np.zeros(16)
    for coded in (False, True):
        for lanes in (1, 2, 4, 8):
            eng = RKEngine(L, "vern7", 1e-8, 1e-10, lanes=lanes, coded=coded)
            eng.set_state(0.0, qd.HipDense.from_array(rho0))
            _lib.call("qa_rk_profile_stages", eng.handle, 3,
                      ms.ctypes.data_as(C.POINTER(C.c_double)))
            _lib.call("qa_device_sync")
    log(f"pmc workload: 3 copies of {COPY_BYTES} B, then 16 x 3 stage launches for each of "
        f"lanes 1, 2, 4, 8, plain then coded (nnz {L.nnz})")


MC_TRAFFIC_FILE = ROOT / "profiles" / "mc_traj_kernel_ising10_traffic.txt"


def read_mc_traffic(prec):
    """HBM bytes of one batched mcsolve launch (FETCH_SIZE + WRITE_SIZE,
    scaled by the copy calibration) from the committed counter summary."""
    try:
        for line in MC_TRAFFIC_FILE.read_text().splitlines():
            k, _, v = line.partition(":")
            if k.strip() == f"traffic_bytes_per_launch_{prec}":
                return int(float(v.strip()))
    except OSError:
        pass
    return None


def mc_probe(reps=2):
    """Only the batched mcsolve launches (TFIM L=10, 2048 trajectories,
    complex128 then complex64), after a 64 MiB device copy used as the
    counter calibration: the command the trajectory kernel's counters are
    read on."""
    from qutip_amd import _lib, data as qd
    from qutip_amd.solver.mcsolve import MCSolver, _read_seeds, draws_for
    from tests import systems
    import qutip_amd as qa  # noqa: F401
    a, b = qd.DeviceBuffer(COPY_BYTES), qd.DeviceBuffer(COPY_BYTES)
    _lib.call("qa_fill_zero", C.c_void_p(a.ptr), C.c_size_t(COPY_BYTES), None)
    for _ in range(3):
        _lib.call("qa_memcpy_d2d", C.c_void_p(b.ptr), C.c_void_p(a.ptr), C.c_size_t(COPY_BYTES),
                  None)
    _lib.call("qa_device_sync")
    s = systems.ising(10, 0.05, tf=20.0, npts=100)
    tlist = np.linspace(0, 20, 100)
    seeds = _read_seeds(list(range(2048)), 2048)
    dev_draws = (qd.DeviceBuffer.from_array(draws_for(seeds, 64)), 2048, 64)
    psi0 = s["psi0"].full().ravel()
    for prec in ("fp64", "fp32"):
        solver = MCSolver(s["H"], s["c_ops"], options={"precision": prec})
        for _ in range(reps):
            solver.run_batch(psi0, tlist, s["e_ops"], dev_draws)
        _lib.call("qa_device_sync")
        print(f"mc probe {prec}: {reps} launches", flush=True)


def jc_probe(reps=6):
    """Only the C1 single solves (Jaynes-Cummings mesolve through
    MESolver.run, one whole-solve launch each): the command the whole-solve
    kernel's counters are read on."""
    import qutip_amd as qa
    from qutip_amd import _lib
    from tests import systems
    s = systems.jaynes_cummings()
    tl = np.load(ROOT / "tests" / "golden" / "jc_mesolve.npz")["tlist"]
    solver = qa.MESolver(s["H"], s["c_ops"], options={"rtol": 1e-8, "atol": 1e-10})
    for _ in range(reps):
        solver.run(s["psi0"], tl, e_ops=s["e_ops"])
    _lib.call("qa_device_sync")
    print(f"jc probe: {reps} solves, {solver._integrator.stats()}", flush=True)


def dense_rhs_probe(N, nc, prec="fp32"):
    """Time the matrix-form Lindblad RHS alone (stage kernels in solve order,
    qa_rk_profile_sequence) for a random N x N system with nc dense collapse
    operators (seeded); prints us per RHS and TFLOP/s."""
    from qutip_amd import _lib
    from qutip_amd.solver.hip_rk import RKEngine
    rng = np.random.default_rng(0)
    H = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    H = H + H.conj().T
    cs = [0.1 * (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))
          for _ in range(nc)]
    Hnh = H - 0.5j * sum(c.conj().T @ c for c in cs)
    eng = RKEngine(None, "vern7", 1e-6, 1e-8, dense=(Hnh, cs), dense_precision=prec)
    eng.set_hermitian(True)  # mesolve's default for a Hermitian rho
    eng.set_state(0.0, (np.eye(N, dtype=complex) / N).ravel(order="F"))
    ts, td = C.c_double(), C.c_double()
    _lib.call("qa_rk_profile_sequence", eng.handle, 10, C.byref(ts), C.byref(td))
    us = ts.value * 1e3
    flop = (1 + 2 * nc) * 8 * N ** 3  # H_nh X, c X, (c X) c^dag: the Hermitian form
    print(json.dumps({"dense_rhs_probe": {"N": N, "nc": nc, "precision": prec,
                                          "hermitian": eng.hermitian,
                                          "us_per_rhs": round(us, 2),
                                          "algorithmic_flop_per_rhs": flop,
                                          "tflops": round(flop / (us * 1e-6) / 1e12, 2)}}),
          flush=True)


def stage_roofline(RKEngine, L, stats, lanes, coded, reps=20):
    """Roofline entry of the stage kernel: HIP events on the engine's stream
    around back-to-back rounds of the 16 stage launches in solve order
    (qa_rk_profile_sequence), weighted by the solve's step and dense-output
    stage counts; algorithmic bytes from stage_bytes."""
    n, nnz = L.shape[0], L.nnz
    eng = RKEngine(L, "vern7", 1e-8, 1e-10, lanes=lanes, coded=coded)
    rho0 = np.zeros(n, dtype=np.complex128)
    rho0[0] = 1.0
    eng.set_state(0.0, rho0)
    form = (2 if eng.sell.uniform else 1) if coded else 0  # the kernel's FORM argument
    t_step, t_dense = C.c_double(), C.c_double()
    _lib_call("qa_rk_profile_sequence", eng.handle, reps, C.byref(t_step), C.byref(t_dense))
    byts = stage_bytes(eng, nnz, n, False)
    n_step = 10 * stats["attempts"]
    n_dense = 6 * stats["dense_preps"]
    avg_ms = (n_step * t_step.value + n_dense * t_dense.value) / (n_step + n_dense)
    avg_bytes = float((stats["attempts"] * byts[:10].sum() + stats["dense_preps"] * byts[10:].sum())
                      / (n_step + n_dense))
    # the algorithmic bytes of one RHS of the CSR complex128 product (values
    # and column indices of every non-zero, the row pointers, x read and
    # y written once); one stage launch evaluates one RHS
    rhs_bytes = nnz * 20 + (n + 1) * 4 + 2 * n * 16
    achieved = rhs_bytes / (avg_ms * 1e-3) / 1e9
    own = avg_bytes / (avg_ms * 1e-3) / 1e9
    log(f"stage kernel ({'coded' if coded else 'plain'}, lanes {lanes}) avg {avg_ms * 1e3:.2f} us "
        f"(step stages {t_step.value * 1e3:.2f}, dense {t_dense.value * 1e3:.2f}): "
        f"{rhs_bytes / 1e6:.2f} MB per RHS -> {achieved:.0f} GB/s; the form's own "
        f"{avg_bytes / 1e6:.2f} MB -> {own:.0f} GB/s")
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": read_traffic(lanes, coded),
            "kernel": f"rk_stage_kernel<double2, {lanes}, {form}>",
            "kernel_form": {0: "plain SELL-C", 1: "coded SELL-C",
                            2: "coded SELL-C, slices of one width"}[form],
            "avg_launch_us": round(avg_ms * 1e3, 3),
            "algorithmic_bytes_per_launch": int(rhs_bytes),
            "algorithmic_unit": "one RHS of the CSR complex128 product: nnz*20 + (n+1)*4 + "
                                "2*n*16 bytes; one RHS per launch",
            "form_bytes_per_launch": int(avg_bytes),
            "form_gbs": round(own, 1),
            "form_bytes_note": "what this operator form and the fused stage combine read and "
                               "write per launch (coded: 2-byte codes and two 256-entry "
                               "tables instead of values and columns)"}


def host_info():
    """The CPU the baselines ran on (lscpu's model and core counts)."""
    info = {"os_cpu_count": os.cpu_count()}
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    try:
        import subprocess
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "CPU(s)", "Thread(s) per core", "Core(s) per socket",
                             "Socket(s)"):
                info[k.strip()] = v.strip()
    except Exception:  # lscpu missing: the counts above remain
        pass
    return info


def _lib_call(*a):
    from qutip_amd import _lib
    _lib.call(*a)


def read_traffic(lanes, coded=False):
    """HBM bytes per stage-kernel launch from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE scaled by the copy calibration) when it was
    taken on the same operator form and SELL shape, else None."""
    vals = {}
    try:
        for line in TRAFFIC_FILE.read_text().splitlines():
            if ":" in line and not line.startswith(("-", "|", "#", "`")):
                k, v = line.split(":", 1)
                vals[k.strip()] = v.strip()
    except OSError:
        return None
    key = f"traffic_bytes_per_launch_{'coded_' if coded else ''}lanes{lanes}"
    return int(float(vals[key])) if key in vals else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-form solve")
    ap.add_argument("--no-mc", action="store_true", help="skip the batched mcsolve leg")
    ap.add_argument("--no-dense", action="store_true", help="skip the dense Kerr leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--serve-requests", type=int, default=2040)
    ap.add_argument("--dense-tf", type=float, default=None,
                    help="end the Kerr solve early (profiling runs; default t=30)")
    ap.add_argument("--dense-reps", type=int, default=2)
    ap.add_argument("--pmc", action="store_true",
                    help="only run the launches the HBM-traffic counters are read on")
    ap.add_argument("--dense-rhs-probe", type=int, nargs=2, metavar=("N", "NC"),
                    help="only time the matrix-form RHS at size N with NC collapse operators")
    ap.add_argument("--mc-probe", action="store_true",
                    help="only run the batched mcsolve launches the trajectory counters are read on")
    ap.add_argument("--jc-probe", action="store_true",
                    help="only run the C1 single solves the whole-solve kernel's counters are read on")
    args = ap.parse_args()
    if args.mc_probe:
        mc_probe()
        return
    if args.jc_probe:
        jc_probe()
        return
    if args.pmc:
        pmc_workload()
        return
    if args.dense_rhs_probe:
        dense_rhs_probe(*args.dense_rhs_probe)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the "
              "solves are independent and would shard one process per GPU, by request "
              "or by trajectory", file=sys.stderr)
        sys.exit(2)

    import qutip_amd as qa
    from qutip_amd import _lib, autotune
    from qutip_amd.solver.hip_rk import RKEngine
    from tests import systems

    log("building the L=8 Ising system")
    s = systems.ising(8, 0.1)
    tlist = s["tlist"]
    opts = {"rtol": 1e-8, "atol": 1e-10}
    solver = qa.MESolver(s["H"], s["c_ops"], options=opts)
    L = solver.rhs.qobj.to_csr()
    n, nnz = L.shape[0], L.nnz

    golden = np.load(ROOT / "tests" / "golden" / "ising8_mesolve.npz")
    log(f"warmup x{args.warmup}")
    for _ in range(args.warmup):
        res = solver.run(s["psi0"], tlist, e_ops=s["e_ops"])
    _lib.call("qa_device_sync")
    log(f"timing x{args.steps}")
    times = []
    t_all = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        res = solver.run(s["psi0"], tlist, e_ops=s["e_ops"])
        _lib.call("qa_device_sync")
        times.append(time.perf_counter() - t0)
    total = time.perf_counter() - t_all
    stats = solver._integrator.stats()
    err = max(np.abs(np.asarray(res.expect[k]) - golden["expect"][k]).max() for k in range(2))
    value = args.steps / total
    log(f"fused: {value:.2f} solves/s, err {err:.2e}, stats {stats}")

    lanes = solver._integrator._engine.lanes
    coded = solver._integrator._engine.coded
    roof = stage_roofline(RKEngine, L, stats, lanes, coded)

    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * total / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "mesolve dissipative TFIM L=8: Liouvillian 65536^2 CSR, 1245183 nnz, "
                        "complex128; vern7 rtol=1e-8 atol=1e-10; t in [0,10], 200 outputs; "
                        "2 e_ops; fused hip_rk (device step control, hipGraph replay)",
            "n": n, "nnz": nnz, "rk_attempts": stats["attempts"],
            "rhs_evals": stats["rhs_evals"], "sell_lanes_per_row": lanes,
            "operator_format": "coded SELL-C" if coded else "SELL-C",
            "autotune": autotune.cache()},
        "err_vs_fp64": float(err),
        "roofline": roof,
    }
    # the plain SELL-C operator (a value and a column per entry) on the same
    # solve: the same result bits, ten times the operator bytes
    ssolver = qa.MESolver(s["H"], s["c_ops"], options={**opts, "operator_format": "sell"})
    sres = ssolver.run(s["psi0"], tlist, e_ops=s["e_ops"])
    _lib.call("qa_device_sync")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sres = ssolver.run(s["psi0"], tlist, e_ops=s["e_ops"])
    _lib.call("qa_device_sync")
    tpl = (time.perf_counter() - t0) / args.steps
    planes = ssolver._integrator._engine.lanes
    diff = max(float(np.abs(np.asarray(sres.expect[k]) - np.asarray(res.expect[k])).max())
               for k in range(2))
    # the same bits hold for the same lane count (the same summation order)
    same = (diff == 0.0) if planes == lanes else None
    out["sell_plain"] = {"solves_per_s": round(1 / tpl, 4), "sell_lanes_per_row": planes,
                         "same_bits_as_coded": same, "max_abs_diff_vs_coded": diff,
                         "roofline": stage_roofline(RKEngine, L, stats, planes, False)}
    log(f"plain SELL-C operator: {1 / tpl:.2f} solves/s (lanes {planes}), max |diff| vs coded "
        f"{diff:.1e}")
    # the chosen precision form of the plain SELL-C operator: complex64
    # values, FP64 state, accumulation, error norm and expectation values
    fsolver = qa.MESolver(s["H"], s["c_ops"], options={**opts, "operator_precision": "fp32",
                                                       "operator_format": "sell"})
    fres = fsolver.run(s["psi0"], tlist, e_ops=s["e_ops"])
    _lib.call("qa_device_sync")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        fres = fsolver.run(s["psi0"], tlist, e_ops=s["e_ops"])
    _lib.call("qa_device_sync")
    tf32 = (time.perf_counter() - t0) / args.steps
    ferr = max(np.abs(np.asarray(fres.expect[k]) - golden["expect"][k]).max() for k in range(2))
    out["fp32_operator"] = {"solves_per_s": round(1 / tf32, 4), "err_vs_fp64": float(ferr),
                            "rk_attempts": fsolver._integrator.stats()["attempts"]}
    log(f"fp32 operator: {1 / tf32:.2f} solves/s, err {ferr:.2e}")
    # the same engine with every iteration's kernels launched directly
    # instead of replaying the captured graph (the form the kernel traces
    # are taken in): the same kernels, so the same result bits
    os.environ["QUTIP_AMD_GRAPH"] = "0"
    try:
        dsolver = qa.MESolver(s["H"], s["c_ops"], options=opts)
    finally:
        del os.environ["QUTIP_AMD_GRAPH"]
    dres = dsolver.run(s["psi0"], tlist, e_ops=s["e_ops"])
    _lib.call("qa_device_sync")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dres = dsolver.run(s["psi0"], tlist, e_ops=s["e_ops"])
    _lib.call("qa_device_sync")
    tdir = (time.perf_counter() - t0) / args.steps
    dsame = all(np.array_equal(np.asarray(dres.expect[k]), np.asarray(res.expect[k]))
                for k in range(2))
    out["direct_launch"] = {"solves_per_s": round(1 / tdir, 4), "same_bits_as_graph": dsame,
                            "graph_launches": dsolver._integrator.stats()["graph_launches"]}
    log(f"direct launches: {1 / tdir:.2f} solves/s, same bits as the graph replay: {dsame}")
    out["jc_mesolve"] = bench_jc(qa, systems, log, cpu=not args.no_cpu)
    if not args.no_mc:
        out["mcsolve"] = bench_mcsolve(qa, systems, log, cpu=not args.no_cpu)
    if not args.no_dense:
        out["dense_kerr"] = bench_dense(qa, systems, log, reps=args.dense_reps, tf=args.dense_tf,
                                        cpu=not args.no_cpu)
    if not args.no_serve:
        out["serving"] = bench_serving(systems, log, nreq=args.serve_requests,
                                       cpu=not args.no_cpu)
    if not args.no_plain:
        psolver = qa.MESolver(s["H"], s["c_ops"], options={**opts, "fused": False})
        t0 = time.perf_counter()
        pres = psolver.run(s["psi0"], tlist, e_ops=s["e_ops"])
        tp = time.perf_counter() - t0
        perr = max(np.abs(np.asarray(pres.expect[k]) - golden["expect"][k]).max()
                   for k in range(2))
        out["plain_form"] = {"solves_per_s": round(1 / tp, 4), "err_vs_fp64": float(perr)}
        log(f"plain: {1 / tp:.3f} solves/s err {perr:.2e}")
    if not args.no_cpu:
        from oracle import baselines, cbind, models
        rho0 = models.vec(models.ket2dm(models.ising(8, 0.1)["psi0"]))
        fun = cbind.functionals_from_ops([O.to_csr() for O in s["e_ops"]], 256)
        log("CPU baseline: C restatement, the full solve on one core, then on every core")
        dt, vals, st = cpu_baseline(L, rho0, tlist, fun)
        cerr = np.abs((vals[:, :2] / vals[:, 2:3]).real.T - golden["expect"]).max()
        allc = baselines.c2()
        out["cpu_baseline"] = {"value": allc["all_cores"], "unit": "solves/s",
                               "cores": allc["cores"], "kind": "port",
                               "one_core": round(1 / dt, 5),
                               "sample": f"{allc['sample']} ({allc['per_process_s']} s each); "
                                         f"one core alone: {dt:.1f} s, {st['n_rhs']} RHS, "
                                         f"err {cerr:.1e}"}
        out["host"] = host_info()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
