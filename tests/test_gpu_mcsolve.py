"""Batched mcsolve on the GPU.

* trajectory by trajectory against the oracle's restatement of MCIntegrator
  with the same seeds: same jump channels, collapse times within the
  secant's norm tolerance, <O>(t) within 1e-6;
* the closed-form jump time of a decaying qubit;
* seed reproducibility (qutip/tests/solver/test_mcsolve.py TestSeeds);
* the dissipative Ising chain L=10 (d=1024, 2048 trajectories, seeds
  0..2047): the ensemble mean within 3 sigma of the FP64 master-equation
  solution, in complex128 and in complex64 with FP64 norms."""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

import qutip_amd as qa  # noqa: E402
from oracle import mc as omc, models, tableaus  # noqa: E402
from tests import systems  # noqa: E402


def _small_ising():
    return systems.ising(4, 0.2, tf=6.0, npts=13)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_trajectories_match_oracle(precision):
    print(f"mc parity vs oracle precision={precision}", flush=True)
    s = _small_ising()
    m = models.ising(4, 0.2, tf=6.0, npts=13)
    ntraj = 24
    res = qa.mcsolve(s["H"], s["psi0"], s["tlist"], s["c_ops"], e_ops=s["e_ops"], ntraj=ntraj,
                     seeds=list(range(ntraj)),
                     options={"keep_runs_results": True, "precision": precision})
    # complex64 norms move the secant's stopping point inside its own band
    # |norm^2 - target| < norm_tol * target: delta t ~ norm_tol / decay rate
    # (the decay rate here is sum_i gamma <n_i> <= 0.8), and <O> after a
    # jump moves with it.
    tol_t = 1e-6 if precision == "fp64" else 1e-3
    tol_e = 1e-6 if precision == "fp64" else 2e-3
    for i in range(ntraj):
        tr = omc.Trajectory(m["H"], m["c_ops"], tableaus.vern7)
        ref = tr.run(m["psi0"], m["tlist"], np.random.default_rng(np.random.SeedSequence(i)),
                     m["e_ops"]).real
        assert [w for _, w in tr.collapses] == list(res.col_which[i]), i
        assert np.allclose([t for t, _ in tr.collapses], res.col_times[i], atol=tol_t, rtol=0), i
        got = np.array([res.runs_expect[k][i] for k in range(2)]).T
        assert np.abs(got - ref).max() < tol_e, (i, np.abs(got - ref).max())


def test_qubit_jump_time_closed_form():
    g = 0.7
    H = qa.Qobj(sp.csr_matrix((2, 2), dtype=complex))
    c = np.sqrt(g) * qa.sigmam()
    tlist = np.linspace(0, 20, 41)
    res = qa.mcsolve(H, qa.basis(2, 0), tlist, [c], e_ops=[qa.sigmaz()], ntraj=16,
                     seeds=list(range(16)), options={"rtol": 1e-10, "atol": 1e-12})
    for i in range(16):
        t_star = -np.log(np.random.default_rng(np.random.SeedSequence(i)).random()) / g
        if t_star < tlist[-1]:
            assert len(res.col_times[i]) == 1
            assert abs(res.col_times[i][0] - t_star) < 2e-4 / g
        else:
            assert len(res.col_times[i]) == 0


def test_seed_reproducibility():
    s = _small_ising()
    a = qa.mcsolve(s["H"], s["psi0"], s["tlist"], s["c_ops"], ntraj=8, seeds=7)
    b = qa.mcsolve(s["H"], s["psi0"], s["tlist"], s["c_ops"], ntraj=8, seeds=7)
    c = qa.mcsolve(s["H"], s["psi0"], s["tlist"], s["c_ops"], ntraj=8, seeds=8)
    assert a.col_times == b.col_times and a.col_which == b.col_which
    assert a.col_times != c.col_times


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_ising10_ensemble_matches_master_equation(golden, precision):
    print(f"ising L=10 mcsolve 2048 traj precision={precision}", flush=True)
    g = golden("ising10_mesolve.npz")
    s = systems.ising(10, 0.05, tf=20.0, npts=100)
    res = qa.mcsolve(s["H"], s["psi0"], g["tlist"], s["c_ops"], e_ops=s["e_ops"], ntraj=2048,
                     seeds=list(range(2048)), options={"precision": precision})
    print(res.stats, flush=True)
    for k in range(2):
        mean = np.asarray(res.average_expect[k])
        sig = np.asarray(res.std_expect[k]) / np.sqrt(2048)
        dev = np.abs(mean - g["expect"][k])
        print(f"e_op {k}: max |mean - ME| = {dev.max():.3e}, max dev/sigma = "
              f"{(dev / np.maximum(sig, 1e-300))[1:].max():.2f}", flush=True)
        assert np.all(dev <= 3 * sig + 1e-8)


def test_plain_drop_in_loop_matches_batched_kernel():
    """MCSolver.run_plain -- qutip's trajectory loop (mcsolve.py:245-406)
    driven through IntegratorHipRK.mcstep and device data ops, as qutip's
    MCSolver runs with method='hip_rk' -- against the batched kernel on the
    same draws: the same jumps and <O>(t) to the step-size rounding."""
    from qutip_amd.solver.mcsolve import _read_seeds, draws_for
    N = 6
    a = qa.destroy(N)
    H = 0.7 * a.dag() @ a + 0.25 * (a + a.dag())
    c_ops = [np.sqrt(0.3) * a, np.sqrt(0.1) * a.dag()]
    e_ops = [a.dag() @ a, a + a.dag()]
    tl = np.linspace(0, 4, 21)
    psi0 = qa.basis(N, 2)
    solver = qa.MCSolver(H, c_ops, options={"rtol": 1e-8, "atol": 1e-10})
    draws = draws_for(_read_seeds(list(range(6)), 6), 64)
    ex, info, _, _ = solver.run_batch(psi0.full().ravel(), tl, e_ops, draws)
    runs, jumps = solver.run_plain(psi0.full().ravel(), tl, e_ops, draws)
    print(f"plain vs batched: jumps {jumps.tolist()} / {info[:, 1].tolist()}, max |d<O>| "
          f"{np.abs(runs - ex).max():.2e}", flush=True)
    assert (info[:, 0] == 0).all()
    assert jumps.tolist() == info[:, 1].tolist()
    assert np.abs(runs - ex).max() < 1e-5
