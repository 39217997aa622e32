"""Whole solves of small systems in one workgroup (qa_rk_small_run) on the
GPU: the same step sequence as the oracle's RK restatement, <O>(t) within
1e-6 of the FP64 matrix exponential, batches of initial states and of
operators, and the failure statuses.

Bar: as the solver tests (1e-6 at rtol=1e-8, atol=1e-10); batched systems
give exactly the single-system result (the same arithmetic in another
workgroup)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from qutip_amd.solver.small import SmallSolve  # noqa: E402
from oracle import cbind, exact, models, tableaus  # noqa: E402

OPTS = dict(rtol=1e-8, atol=1e-10)


def _jc(g=0.05 * 2 * np.pi):
    m = models.jaynes_cummings(g=g)
    Lm = models.liouvillian(m["H"], m["c_ops"])
    return m, Lm


def _functionals(e_ops, d):
    fptr, fidx, fw = cbind.functionals_from_ops(e_ops, d)
    kinds = np.zeros(len(fptr) - 1, np.int32)
    return (kinds, fptr.astype(np.int32), np.zeros(len(fidx), np.int32), fidx.astype(np.int32),
            fw.view(np.complex128))


def _expect(vals):
    """<O> = functional / trace functional (the last one)."""
    return (vals[..., :-1] / vals[..., -1:]).real


def test_jc_whole_solve_matches_oracle_and_exact():
    print("small: JC whole solve", flush=True)
    m, Lm = _jc()
    tl = m["tlist"]
    ss = SmallSolve(Lm, "vern7", **OPTS)
    ss.set_functionals(*_functionals(m["e_ops"], 20))
    rho0 = models.vec(models.ket2dm(m["psi0"]))
    vals, yfin, info = ss.run(tl[0], rho0, tl[1:])
    assert info[0, 0] >= 0 and info[0, 1] == len(tl) - 1
    _, _, ost = cbind.rk_solve(Lm, rho0, tl, tableaus.vern7,
                               cbind.functionals_from_ops(m["e_ops"], 20), **OPTS)
    print(f"  attempts {info[0, 3]} rhs {ss.rhs_evals(info[0])}; oracle {ost}", flush=True)
    assert info[0, 3] == ost["n_attempts"]
    assert ss.rhs_evals(info[0]) == ost["n_rhs"]
    ref = exact.mesolve_expect(m["H"], m["c_ops"], m["psi0"], tl, m["e_ops"])
    err = np.abs(_expect(vals[0]).T - ref[:, 1:]).max()
    print(f"  max err vs expm {err:.2e}", flush=True)
    assert err < 1e-6
    # final state = the state at the last output time
    yf = np.zeros(400, complex)
    yfin.download(yf)
    v_last = exact.states_uniform(Lm, rho0, tl)[-1]
    assert np.abs(yf - v_last).max() < 1e-6


def test_batch_of_states_equals_single_runs():
    print("small: batch of 64 initial states", flush=True)
    m, Lm = _jc()
    tl = np.linspace(0, 5, 21)
    rng = np.random.default_rng(3)
    psis = rng.standard_normal((64, 20)) + 1j * rng.standard_normal((64, 20))
    psis /= np.linalg.norm(psis, axis=1, keepdims=True)
    rhos = np.array([models.vec(np.outer(p, p.conj())) for p in psis])
    ss = SmallSolve(Lm, "vern7", **OPTS)
    ss.set_functionals(*_functionals(m["e_ops"], 20))
    vals, _, info = ss.run(0.0, rhos, tl[1:])
    assert (info[:, 0] >= 0).all()
    for i in (0, 17, 63):
        one, _, _ = ss.run(0.0, rhos[i], tl[1:])
        assert np.array_equal(one[0], vals[i]), i
        ref = exact.mesolve_expect(m["H"], m["c_ops"], np.outer(psis[i], psis[i].conj()), tl,
                                   m["e_ops"])
        assert np.abs(_expect(vals[i]).T - ref[:, 1:]).max() < 1e-6, i


def test_batch_of_operators():
    print("small: 8 couplings in one launch", flush=True)
    gs = 2 * np.pi * np.linspace(0.02, 0.09, 8)
    Ls, ms = [], []
    for g in gs:
        m, Lm = _jc(g)
        ms.append(m)
        Ls.append(Lm)
    tl = np.linspace(0, 10, 41)
    ss = SmallSolve(Ls, "vern7", **OPTS)
    ss.set_functionals(*_functionals(ms[0]["e_ops"], 20))
    rho0 = models.vec(models.ket2dm(ms[0]["psi0"]))
    vals, _, info = ss.run(0.0, np.tile(rho0, (8, 1)), tl[1:])
    assert (info[:, 0] >= 0).all()
    for i, m in enumerate(ms):
        ref = exact.mesolve_expect(m["H"], m["c_ops"], m["psi0"], tl, m["e_ops"])
        err = np.abs(_expect(vals[i]).T - ref[:, 1:]).max()
        assert err < 1e-6, (i, err)


def test_sesolve_quadratic_functionals():
    print("small: sesolve ket with quadratic functionals", flush=True)
    L = 6
    m = models.ising(L, 0.1)
    H = m["H"]
    d = H.shape[0]
    tl = np.linspace(0, 3, 31)
    ss = SmallSolve(-1j * H, "vern7", **OPTS)
    kinds, fptr, rows, cols, w = [], [0], [], [], []
    import scipy.sparse as sp
    for O in m["e_ops"] + [sp.identity(d, format="csr")]:
        C = sp.coo_matrix(O)
        kinds.append(1)
        rows.append(C.row)
        cols.append(C.col)
        w.append(C.data.astype(complex))
        fptr.append(fptr[-1] + C.nnz)
    ss.set_functionals(np.array(kinds, np.int32), np.array(fptr, np.int32),
                       np.concatenate(rows).astype(np.int32),
                       np.concatenate(cols).astype(np.int32), np.concatenate(w))
    vals, _, info = ss.run(0.0, m["psi0"], tl[1:])
    assert info[0, 0] >= 0
    ref = exact.sesolve_expect(H, m["psi0"], tl, m["e_ops"])
    assert np.abs(_expect(vals[0]).T - ref[:, 1:]).max() < 1e-6


def test_statuses():
    print("small: too much work, dt underflow", flush=True)
    m, Lm = _jc()
    rho0 = models.vec(models.ket2dm(m["psi0"]))
    ss = SmallSolve(Lm, "vern7", nsteps=2, **OPTS)
    ss.set_functionals(*_functionals(m["e_ops"], 20))
    _, _, info = ss.run(0.0, rho0, [25.0])
    assert info[0, 0] == -1  # TOO_MUCH_WORK
    # a step pinned at 0.5 cannot meet rtol 1e-10 (the oracle's restatement
    # stops with DT_UNDERFLOW after one attempt on the same system)
    ss = SmallSolve(Lm, "vern7", min_step=0.5, max_step=0.5, rtol=1e-10, atol=1e-12)
    _, _, info = ss.run(0.0, rho0, [2.0])
    _, _, ost = cbind.rk_solve(Lm, rho0, [0.0, 2.0], tableaus.vern7,
                               cbind.functionals_from_ops(m["e_ops"], 20), rtol=1e-10,
                               atol=1e-12, min_step=0.5, max_step=0.5)
    assert ost["status"] == -2
    assert info[0, 0] == -2 and info[0, 3] == ost["n_attempts"]  # DT_UNDERFLOW


def test_solver_takes_the_initial_point_from_the_kernel():
    """MESolver.run on the whole-solve path evaluates the e_ops at tlist[0]
    in the kernel (the functionals of the initial state): the values equal
    tr(O rho0) computed on the host, and a second run with the same e_ops
    list (functionals kept from the first run) gives the same bits; a new
    list with other operators gives its own values."""
    print("small: solver initial point from the kernel", flush=True)
    import qutip_amd as qa
    from tests import systems
    import pathlib
    s = systems.jaynes_cummings()
    g = np.load(pathlib.Path(__file__).parent / "golden" / "jc_mesolve.npz")
    tl = g["tlist"]
    solver = qa.MESolver(s["H"], s["c_ops"], options=OPTS)
    r1 = solver.run(s["psi0"], tl, e_ops=s["e_ops"])
    assert solver._integrator.stats()["path"] == "whole solve in one workgroup"
    assert solver._integrator.stats()["outputs"] == len(tl)
    psi = s["psi0"].full().ravel()
    for k, O in enumerate(s["e_ops"]):
        want = np.vdot(psi, O.full() @ psi).real
        assert len(r1.expect[k]) == len(tl)
        assert abs(r1.expect[k][0] - want) < 1e-14, (k, r1.expect[k][0], want)
    r2 = solver.run(s["psi0"], tl, e_ops=s["e_ops"])
    for k in range(2):
        assert np.array_equal(np.asarray(r1.expect[k]), np.asarray(r2.expect[k]))
    # other operators: the kept functionals must not be reused
    ops3 = [s["e_ops"][1], s["e_ops"][0]]
    r3 = solver.run(s["psi0"], tl, e_ops=ops3)
    assert np.array_equal(np.asarray(r3.expect[0]), np.asarray(r1.expect[1]))
    assert np.array_equal(np.asarray(r3.expect[1]), np.asarray(r1.expect[0]))
    for k in range(2):  # the golden QuTiP run on the same times
        assert np.abs(np.asarray(r1.expect[k]) - g["expect"][k]).max() < 1e-6


def test_sesolve_initial_point_from_the_kernel():
    """SESolver.run on the whole-solve path (a ket: quadratic functionals
    conj(psi) O psi) takes the e_ops values at tlist[0] from the kernel:
    they equal <psi0|O|psi0> on the host, and every value is within 1e-6 of
    the FP64 matrix exponential."""
    print("small: sesolve initial point from the kernel", flush=True)
    import qutip_amd as qa
    from tests import systems
    s = systems.jaynes_cummings()
    psi0 = (qa.tensor(qa.basis(10, 2), qa.basis(2, 0)) +
            0.5j * qa.tensor(qa.basis(10, 1), qa.basis(2, 1))).unit()
    tl = np.linspace(0.0, 5.0, 41)
    solver = qa.SESolver(s["H"], options=OPTS)
    r = solver.run(psi0, tl, e_ops=s["e_ops"])
    assert solver._integrator.stats()["path"] == "whole solve in one workgroup"
    psi = psi0.full().ravel()
    ref = exact.sesolve_expect(s["H"].full(), psi, tl, [O.full() for O in s["e_ops"]])
    for k, O in enumerate(s["e_ops"]):
        want = np.vdot(psi, O.full() @ psi).real
        got = np.asarray(r.expect[k])
        assert abs(got[0] - want) < 1e-14, (k, got[0], want)
        assert np.abs(got - ref[k]).max() < 1e-6, k
