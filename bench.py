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
    """Algorithmic HBM bytes of each stage-kernel launch: operator values and
    columns once, the input vector once, k_i written, plus the epilogue's
    reads of y_prev and of every k_j with a non-zero coefficient and its
    write of the next stage input (or y_front)."""
    tab = engine.tableau
    a, b, e = tab["a"], tab["b"], tab["e"]
    s, se = len(b), len(tab["c"])
    vb = 8 if fp32 else 16
    base = nnz * (vb + 4) + (engine.sell.nslices + 1) * 4 + n * 16 + n * 16
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


def bench_mcsolve(qa, systems, log, ntraj=2048, reps=3):
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
        out[prec] = {"traj_per_s": round(ntraj / best, 1), "s_per_batch": round(best, 4),
                     "all_ok": ok, "jumps": int(info[:, 1].sum()),
                     "rk_attempts": int(info[:, 3].sum()),
                     "max_dev_over_sigma_vs_mesolve": round(worst_sigma, 3)}
        log(f"mcsolve {prec}: {ntraj / best:.0f} traj/s ({best * 1e3:.1f} ms), worst "
            f"{worst_sigma:.2f} sigma, jumps {int(info[:, 1].sum())}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-form solve")
    ap.add_argument("--no-mc", action="store_true", help="skip the batched mcsolve leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the "
              "solves are independent and would shard by request/trajectory, see DESIGN.md",
              file=sys.stderr)
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

    # roofline of the stage kernel (HIP events on the engine's stream)
    lanes = solver._integrator._engine.lanes
    eng = RKEngine(L, "vern7", 1e-8, 1e-10, lanes=lanes)
    ms = np.zeros(16)
    _lib.call("qa_rk_profile_stages", eng.handle, 20, ms.ctypes.data_as(C.POINTER(C.c_double)))
    byts = stage_bytes(eng, nnz, n, False)
    cnt = np.zeros(16)
    cnt[:10] = stats["attempts"]
    cnt[10:] = stats["dense_preps"]
    avg_ms = float((cnt * ms).sum() / cnt.sum())
    avg_bytes = float((cnt * byts).sum() / cnt.sum())
    achieved = avg_bytes / (avg_ms * 1e-3) / 1e9
    log(f"stage kernel avg {avg_ms * 1e3:.2f} us, {avg_bytes / 1e6:.2f} MB -> {achieved:.0f} GB/s;"
        f" per stage us {np.round(ms * 1e3, 2).tolist()}")

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
            "autotune": autotune.cache()},
        "err_vs_fp64": float(err),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": f"rk_stage_kernel<double2, {lanes}>",
                     "avg_launch_us": round(avg_ms * 1e3, 3),
                     "algorithmic_bytes_per_launch": int(avg_bytes)},
    }
    if not args.no_mc:
        out["mcsolve"] = bench_mcsolve(qa, systems, log)
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
        from oracle import cbind, models
        rho0 = models.vec(models.ket2dm(models.ising(8, 0.1)["psi0"]))
        fun = cbind.functionals_from_ops([O.to_csr() for O in s["e_ops"]], 256)
        log("CPU baseline: C restatement, 1 core, the full solve")
        dt, vals, st = cpu_baseline(L, rho0, tlist, fun)
        cerr = np.abs((vals[:, :2] / vals[:, 2:3]).real.T - golden["expect"]).max()
        out["cpu_baseline"] = {"value": round(1 / dt, 5), "unit": "solves/s", "cores": 1,
                               "kind": "port",
                               "sample": f"one full solve of the same workload ({dt:.1f} s, "
                                         f"{st['n_rhs']} RHS, err {cerr:.1e})"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
