"""Time-dependent generators (QuTiP's list form) on the GPU.

* the coefficient programs evaluated by the device (qa_coef_eval) against
  the host coefficients (string coefficients evaluated as the reference's
  StrFunctionCoefficient does, interpolated arrays as InterCoefficient);
* sesolve / mesolve of a qubit driven by a Gaussian pulse f(t) sigma_x on
  the fused engine (coefficients evaluated by the stage kernels) and on the
  plain form, against the closed forms cos 2F(t) and cos 2F(t) e^{-2 g t};
* a driven TFIM chain (string and array coefficients together) on the
  fused engine and the plain form against the oracle's RK on the same
  time-dependent right-hand side."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

import qutip_amd as qa  # noqa: E402
from qutip_amd import _lib, coefficient as qc, data as qd  # noqa: E402
from oracle import models, rk, tableaus  # noqa: E402
from tests import systems  # noqa: E402
from tests.test_coefficient import EXPRESSIONS  # noqa: E402

ARGS = {"A": 2.0, "t0": 1.5, "s": 0.4}
PULSE = "A * exp(-(t - t0) ** 2 / (2 * s ** 2))"


def _F(t):
    from scipy.special import erf
    A, t0, s = ARGS["A"], ARGS["t0"], ARGS["s"]
    c = s * np.sqrt(2)
    return A * s * np.sqrt(np.pi / 2) * (erf((t - t0) / c) + erf(t0 / c))


def test_device_coefficients_match_host():
    cs = [qc.coefficient(e, args=a) for e, a in EXPRESSIONS]
    tl = np.linspace(-0.5, 3.7, 9)
    cs += [qc.coefficient(np.cos(tl) + 1j * tl, tlist=tl, order=k) for k in (0, 1, 3)]
    times = np.linspace(-0.3, 4.0, 37)
    tdev = qd.DeviceBuffer.from_array(times)
    for g0 in range(0, len(cs), 8):
        group = cs[g0:g0 + 8]  # at most 8 programs per set
        start, code, consts = qc.DeviceProgram.concat([c.device_program() for c in group]).arrays()
        out = np.zeros(len(group) * len(times), np.complex128)
        O = qd.DeviceBuffer(out.nbytes)
        _lib.call("qa_coef_eval", len(group), start.ctypes.data_as(C.POINTER(C.c_int)),
                  len(code), code.ctypes.data_as(C.POINTER(C.c_int)), len(consts),
                  consts.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)), len(times),
                  C.c_void_p(tdev.ptr), C.c_void_p(O.ptr), None)
        O.download(out)
        got = out.reshape(len(group), len(times))
        for k, c in enumerate(group):
            want = np.array([c(t) for t in times])
            err = np.abs(got[k] - want) / np.maximum(1.0, np.abs(want))
            print(f"coefficient {k}: max rel err {err.max():.1e}", flush=True)
            assert err.max() < 1e-13


@pytest.mark.parametrize("fused", [True, False])
def test_driven_qubit_sesolve_closed_form(fused):
    print(f"driven qubit sesolve fused={fused}", flush=True)
    tl = np.linspace(0, 3, 31)
    H = [[qa.sigmax(), PULSE]]
    solver = qa.SESolver(qa.QobjEvo(H, args=ARGS),
                         options={"rtol": 1e-8, "atol": 1e-10, "fused": fused})
    res = solver.run(qa.basis(2, 0), tl, e_ops=[qa.sigmaz()])
    name = solver._integrator.name
    err = np.abs(np.asarray(res.expect[0]) - np.cos(2 * _F(tl))).max()
    print(f"  {name}: max err {err:.2e}", flush=True)
    assert ("time-dependent" in name) == fused
    assert err < 1e-6


@pytest.mark.parametrize("fused", [True, False])
def test_driven_qubit_mesolve_closed_form(fused):
    print(f"driven qubit mesolve fused={fused}", flush=True)
    g = 0.3
    tl = np.linspace(0, 3, 31)
    res = qa.mesolve([[qa.sigmax(), PULSE]], qa.basis(2, 0), tl, [np.sqrt(g) * qa.sigmax()],
                     e_ops=[qa.sigmaz()], args=ARGS,
                     options={"rtol": 1e-8, "atol": 1e-10, "fused": fused})
    want = np.cos(2 * _F(tl)) * np.exp(-2 * g * tl)
    err = np.abs(np.asarray(res.expect[0]) - want).max()
    print(f"  {res.stats.get('method')}: max err {err:.2e}", flush=True)
    assert err < 1e-6


def _driven_chain(L):
    s = systems.ising(L, 0.1, tf=4.0, npts=41)
    xs = systems.site(qa.sigmax(), 0, L)
    for i in range(1, L):
        xs = xs + systems.site(qa.sigmax(), i, L)
    zs = systems.site(qa.sigmaz(), 0, L)
    tl_c = np.linspace(0, 4, 17)
    H = qa.QobjEvo([s["H"], [xs, "0.5 * cos(w * t)"], [zs, 0.3 * np.sin(tl_c)]],
                   args={"w": 1.7}, tlist=tl_c)
    return s, H


@pytest.mark.parametrize("fused", [True, False])
def test_driven_chain_mesolve_vs_oracle(fused):
    """TFIM L=4 mesolve with a cosine drive on every sigma_x and an
    interpolated field on sigma_z of site 0; the oracle integrates the same
    list-form right-hand side with the host coefficients."""
    L = 4
    print(f"driven TFIM L={L} mesolve fused={fused}", flush=True)
    s, H = _driven_chain(L)
    opts = {"rtol": 1e-8, "atol": 1e-10, "fused": fused}
    solver = qa.MESolver(H, s["c_ops"], options=opts)
    res = solver.run(s["psi0"], s["tlist"], e_ops=s["e_ops"])
    Lt = solver.rhs
    parts = [(q.to_csr(), c) for q, c in Lt.elements]
    f = models.td_rhs(Lt.qobj.to_csr(), parts)
    rho0 = models.ket2dm(s["psi0"].full().ravel())
    ys, ork = rk.solve(f, models.vec(rho0), s["tlist"], tableaus.TABLEAUS["vern7"],
                       rtol=1e-10, atol=1e-12)
    ops = [O.to_csr() for O in s["e_ops"]]
    d = 2 ** L
    ref = np.array([[models.expect_dm(O, models.unvec(y, d)).real for y in ys] for O in ops])
    err = max(np.abs(np.asarray(res.expect[k]) - ref[k]).max() for k in range(2))
    print(f"  {solver._integrator.name}: max err {err:.2e}; stats {res.stats}", flush=True)
    assert ("time-dependent" in solver._integrator.name) == fused
    assert err < 1e-6


def test_fused_time_dependent_follows_the_oracle_step_sequence():
    """Same tolerances as the oracle: the fused engine takes the oracle's
    step sequence on a driven ket (the coefficients enter each stage at
    t_prev + c_i dt, explicit_rk.pyx:358-389)."""
    L = 6
    s, H = _driven_chain(L)
    opts = {"rtol": 1e-8, "atol": 1e-10}
    solver = qa.SESolver(H, options=opts)
    res = solver.run(s["psi0"], s["tlist"], e_ops=s["e_ops"])
    A = solver.rhs
    f = models.td_rhs(A.qobj.to_csr(), [(q.to_csr(), c) for q, c in A.elements])
    ys, ork = rk.solve(f, s["psi0"].full().ravel(), s["tlist"], tableaus.TABLEAUS["vern7"],
                       rtol=1e-8, atol=1e-10)
    st = solver._integrator._engine.stats()
    print(f"fused attempts {st['attempts']} vs oracle {ork.n_attempts}", flush=True)
    assert abs(st["attempts"] - ork.n_attempts) <= 2
    ops = [O.to_csr() for O in s["e_ops"]]
    ref = np.array([[models.expect_ket(O, y).real / np.vdot(y, y).real for y in ys] for O in ops])
    err = max(np.abs(np.asarray(res.expect[k]) - ref[k]).max() for k in range(2))
    assert err < 1e-6
