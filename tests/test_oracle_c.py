"""The C oracle (CPU baseline) against the reference's own C++ spmv
(oracle/_ref, built from qutip/core/data/src/matmul_csr_vector.cpp) and
against the numpy restatement of explicit_rk.pyx (CPU only)."""
import numpy as np
import pytest
import scipy.sparse as sp
from numpy.testing import assert_allclose

from oracle import cbind, models, rk, tableaus

needs_ref = pytest.mark.skipif(not cbind.ref_available(),
                               reason="oracle/_ref not built (reference sources absent)")


def _rand_csr(n, m, density, seed):
    rng = np.random.default_rng(seed)
    A = sp.random(n, m, density=density, random_state=rng, format="csr", dtype=np.float64)
    A = A + 1j * sp.random(n, m, density=density, random_state=rng, format="csr")
    return sp.csr_matrix(A)


@needs_ref
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_c_spmv_bit_exact_vs_reference(seed):
    A = _rand_csr(300, 250, 0.05, seed)
    rng = np.random.default_rng(seed + 10)
    x = rng.standard_normal(250) + 1j * rng.standard_normal(250)
    y0 = rng.standard_normal(300) + 1j * rng.standard_normal(300)
    a = cbind.csr_matvec(A, x, y0.copy(), scale=0.3 - 1.1j)
    b = cbind.ref_csr_matvec(A, x, y0.copy(), scale=0.3 - 1.1j)
    assert np.array_equal(a, b)
    assert_allclose(a, y0 + (0.3 - 1.1j) * (A @ x), rtol=1e-12, atol=1e-12)


@needs_ref
def test_c_spmv_bit_exact_on_benchmark_liouvillian():
    m = models.jaynes_cummings()
    Lm = models.liouvillian(m["H"], m["c_ops"])
    x = np.random.default_rng(5).standard_normal(400) + 0j
    assert np.array_equal(cbind.csr_matvec(Lm, x), cbind.ref_csr_matvec(Lm, x))


def test_c_spmv_empty_rows_and_zero_matrix():
    A = sp.csr_matrix((7, 7), dtype=np.complex128)
    x = np.ones(7, dtype=np.complex128)
    assert np.array_equal(cbind.csr_matvec(A, x), np.zeros(7))


@pytest.mark.parametrize("tab", ["vern7", "tsit5"])
def test_c_rk_matches_numpy_rk(tab):
    m = models.jaynes_cummings()
    Lm = models.liouvillian(m["H"], m["c_ops"])
    rho0 = models.vec(models.ket2dm(m["psi0"]))
    fun = cbind.functionals_from_ops(m["e_ops"], 20)
    T = tableaus.TABLEAUS[tab]
    vals, yf, st = cbind.rk_solve(Lm, rho0, m["tlist"], T, fun, rtol=1e-8, atol=1e-10)
    ys, r = rk.solve(lambda t, y: Lm @ y, rho0, m["tlist"], T, rtol=1e-8, atol=1e-10)
    assert st["n_rhs"] == r.n_rhs and st["n_attempts"] == r.n_attempts
    assert_allclose(yf, ys[-1], rtol=1e-13, atol=1e-15)


def test_c_rk_matches_golden_jc(golden):
    g = golden("jc_mesolve.npz")
    m = models.jaynes_cummings()
    Lm = models.liouvillian(m["H"], m["c_ops"])
    fun = cbind.functionals_from_ops(m["e_ops"], 20)
    vals, _, st = cbind.rk_solve(Lm, models.vec(models.ket2dm(m["psi0"])), g["tlist"],
                                 tableaus.vern7, fun, rtol=1e-8, atol=1e-10)
    e = (vals[:, :2] / vals[:, 2:3]).real.T
    assert np.abs(e - g["expect"]).max() < 1e-7
    assert st["status"] >= 0


def test_c_rk_too_much_work_status():
    m = models.jaynes_cummings()
    Lm = models.liouvillian(m["H"], m["c_ops"])
    fun = cbind.functionals_from_ops(m["e_ops"], 20)
    _, _, st = cbind.rk_solve(Lm, models.vec(models.ket2dm(m["psi0"])), [0.0, 25.0],
                              tableaus.vern7, fun, rtol=1e-8, atol=1e-10, nsteps=2)
    assert st["status"] == -1


def test_c_mc_trajectories_match_the_numpy_oracle():
    """The C trajectory loop (the C3 CPU baseline) against oracle/mc.py on
    the same draws: the same jumps and <O>(t) within 1e-8."""
    from oracle import mc as omc, models, tableaus
    from oracle.cbind import McTrajectories
    m = models.ising(4, 0.2, tf=4.0, npts=9)
    cm = McTrajectories(m["H"], m["c_ops"], m["e_ops"], tableaus.TABLEAUS["vern7"])
    for seed in range(6):
        rng = np.random.default_rng(seed)
        draws = np.random.default_rng(seed).random(64)
        tr = omc.Trajectory(m["H"], m["c_ops"], tableaus.TABLEAUS["vern7"])
        want = tr.run(m["psi0"], m["tlist"], rng, m["e_ops"])
        got, st = cm.run(m["psi0"], m["tlist"], draws)
        assert st["rc"] == 0 and st["jumps"] == len(tr.collapses)
        assert st["draws"] == tr.draws_used
        assert np.abs(got - want).max() < 1e-8


def test_c_banded_matrix_form_vs_exact():
    """The C banded matrix-form RK (the C4 CPU baseline) on a small driven
    Kerr oscillator against the FP64 matrix exponential of the Liouvillian."""
    from oracle import cbind, exact, models, tableaus
    m = models.kerr(16)
    tl = np.linspace(0, 3, 7)
    rho0 = models.vec(models.ket2dm(m["psi0"]))
    fun = cbind.functionals_from_ops(m["e_ops"], 16)
    vals, st = cbind.rk_solve_dia(m["H"], m["c_ops"], rho0, tl, tableaus.TABLEAUS["vern7"], fun)
    got = (vals[:, :2] / vals[:, 2:3]).real.T
    ref = exact.mesolve_expect(m["H"], m["c_ops"], m["psi0"], tl, m["e_ops"])
    assert st["status"] >= 0 and np.abs(got - ref).max() < 1e-7
