"""The coded SELL-C operator (2-byte codes into value and column-offset
tables) on the GPU: the same bits as the plain SELL-C form, product by
product and solve by solve, and the oracle's step sequence.

Bar: bit-identical to the plain form (the same FP64 terms summed in the same
order); the solve's attempts and RHS count equal the oracle's restatement of
explicit_rk.pyx and its expectation values are within 1e-6 of the FP64
matrix exponential (rtol=1e-8, atol=1e-10)."""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

from qutip_amd import data as qd  # noqa: E402
from qutip_amd.solver.hip_rk import RKEngine  # noqa: E402
from oracle import cbind, exact, models, tableaus  # noqa: E402


def _operators():
    m = models.ising(6, 0.1)
    Lm = models.liouvillian(m["H"], m["c_ops"]).tocsr()
    rng = np.random.default_rng(5)
    n = 3001  # ragged: not a multiple of any slice height
    band = sp.diags([rng.choice([0.5, -1j, 2.0 + 0.25j], n - abs(k)) for k in (0, 1, -7, 130)],
                    [0, 1, -7, 130], shape=(n, n), format="csr")
    return {"ising6_liouvillian": Lm, "banded_3001": band}


@pytest.mark.parametrize("lanes", [1, 2, 4, 8])
@pytest.mark.parametrize("name", ["ising6_liouvillian", "banded_3001"])
def test_coded_matvec_equals_plain(name, lanes):
    print(f"coded matvec {name} lanes={lanes}", flush=True)
    A = _operators()[name]
    n = A.shape[0]
    rng = np.random.default_rng(lanes)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    xd = qd.HipDense.from_array(x)
    outs = []
    for coded in (False, True):
        M = qd.SellMatrix(A, lanes=lanes, coded=coded)
        out = qd.HipDense((n, 1))
        M.matvec(xd, out)
        outs.append(out.to_array().ravel())
    assert np.array_equal(outs[0], outs[1])
    np.testing.assert_allclose(outs[1], A @ x, rtol=0, atol=1e-12 * np.abs(A).sum(axis=1).max())


@pytest.mark.parametrize("lanes", [1, 4])
def test_coded_solve_equals_plain_and_oracle(lanes):
    """Dissipative Ising L=5 Liouvillian (1024 x 1024) through the fused
    engine: plain, coded with per-slice widths and coded with one width for
    all slices give the same functional and state bits, the oracle's attempt
    count, and <O>(t) within 1e-6 of expm."""
    print(f"coded solve Ising L=5 lanes={lanes}", flush=True)
    m = models.ising(5, 0.1)
    Lm = models.liouvillian(m["H"], m["c_ops"]).tocsr()
    d = m["H"].shape[0]
    rho0 = models.vec(models.ket2dm(m["psi0"]))
    tl = np.linspace(0.0, 2.0, 21)
    fptr, fidx, fw = cbind.functionals_from_ops(m["e_ops"], d)
    res = {}
    for form in ("plain", "sliced", "uniform"):
        eng = RKEngine(Lm, "vern7", 1e-8, 1e-10, lanes=lanes, coded=form != "plain",
                       uniform=form == "uniform")
        if form == "uniform":
            assert eng.sell.uniform
        if form == "sliced" and lanes == 1:  # (at 4 lanes every slice has one width anyway)
            assert not eng.sell.uniform
        eng.set_functionals(np.zeros(len(fptr) - 1, np.int32), fptr.astype(np.int32),
                            np.zeros(len(fidx), np.int32), fidx.astype(np.int32),
                            fw.view(np.complex128))
        eng.set_state(tl[0], rho0)
        vals, st, _ = eng.run(tl[1:])
        assert st >= 0
        _, y = eng.get_state()
        res[form] = (vals, y.to_array().ravel(), eng.stats())
    for form in ("sliced", "uniform"):
        assert np.array_equal(res["plain"][0], res[form][0]), form
        assert np.array_equal(res["plain"][1], res[form][1]), form
        assert res["plain"][2]["attempts"] == res[form][2]["attempts"], form
    _, _, ost = cbind.rk_solve(Lm, rho0, tl, tableaus.vern7, (fptr, fidx, fw), rtol=1e-8,
                               atol=1e-10)
    assert res["uniform"][2]["attempts"] == ost["n_attempts"]
    ref = exact.mesolve_expect(m["H"], m["c_ops"], m["psi0"], tl, m["e_ops"])
    vals = res["uniform"][0]
    got = (vals[:, :-1] / vals[:, -1:]).real
    assert np.abs(got.T - ref[:, 1:]).max() < 1e-6


def test_coded_fp32_equals_plain_fp32():
    """operator_precision fp32: the coded table holds the complex64-rounded
    values, so the products equal the plain complex64 form's."""
    print("coded matvec fp32", flush=True)
    A = _operators()["ising6_liouvillian"]
    n = A.shape[0]
    x = np.random.default_rng(9).standard_normal(n) + 0j
    xd = qd.HipDense.from_array(x)
    outs = []
    for coded in (False, True):
        M = qd.SellMatrix(A, fp32=True, lanes=2, coded=coded)
        out = qd.HipDense((n, 1))
        M.matvec(xd, out)
        outs.append(out.to_array().ravel())
    assert np.array_equal(outs[0], outs[1])


def test_mesolve_default_uses_the_coded_operator(monkeypatch):
    """MESolver's fused integrator (Ising L=6: Liouvillian 4096 x 4096, above
    the whole-solve size) picks the coded form for an operator that fits the
    tables, with the plain form's expectation values bit for bit (the same
    lane count for both: the heuristic one, no timing)."""
    print("mesolve auto coded", flush=True)
    monkeypatch.setenv("QUTIP_AMD_NO_TUNE", "1")
    import qutip_amd as qa
    from tests import systems
    s = systems.ising(6, 0.1, tf=2.0, npts=21)
    solver = qa.MESolver(s["H"], s["c_ops"], options={"rtol": 1e-8, "atol": 1e-10})
    res = solver.run(s["psi0"], s["tlist"], e_ops=s["e_ops"])
    assert solver._integrator._engine.coded
    assert "coded" in solver._integrator.name
    sref = qa.MESolver(s["H"], s["c_ops"], options={"rtol": 1e-8, "atol": 1e-10,
                                                    "operator_format": "sell"})
    ref = sref.run(s["psi0"], s["tlist"], e_ops=s["e_ops"])
    assert not sref._integrator._engine.coded
    assert sref._integrator._engine.lanes == solver._integrator._engine.lanes
    for k in range(len(s["e_ops"])):
        assert np.array_equal(np.asarray(res.expect[k]), np.asarray(ref.expect[k]))


def test_sesolve_fused_coded_quadratic_functionals(monkeypatch):
    """sesolve of TFIM L=11 (ket of 2048 entries, above the whole-solve size):
    the fused engine with the coded operator and quadratic e_ops
    functionals, within 1e-6 of the FP64 matrix exponential and bit for bit
    the plain form's values at the same lane count."""
    print("sesolve L=11 fused coded", flush=True)
    import qutip_amd as qa
    from tests import systems
    monkeypatch.setenv("QUTIP_AMD_NO_TUNE", "1")
    s = systems.ising(11, 0.1, tf=2.0, npts=11)
    m = models.ising(11, 0.1)
    opts = {"rtol": 1e-8, "atol": 1e-10}
    solver = qa.SESolver(s["H"], options=opts)
    res = solver.run(s["psi0"], s["tlist"], e_ops=s["e_ops"])
    assert solver._integrator._engine.coded
    ref = exact.sesolve_expect(m["H"], m["psi0"], s["tlist"], m["e_ops"])
    err = max(np.abs(np.asarray(res.expect[k]) - ref[k]).max() for k in range(2))
    print(f"  max err vs expm {err:.2e}", flush=True)
    assert err < 1e-6
    plain = qa.SESolver(s["H"], options={**opts, "operator_format": "sell"})
    pres = plain.run(s["psi0"], s["tlist"], e_ops=s["e_ops"])
    assert not plain._integrator._engine.coded
    for k in range(2):
        assert np.array_equal(np.asarray(res.expect[k]), np.asarray(pres.expect[k]))
