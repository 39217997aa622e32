"""sesolve / mesolve through hip_rk on the GPU, against the FP64 golden
fixtures (matrix exponential) and the oracle's RK restatement.

Bar: <O>(t) within 1e-6 of the FP64 solution at rtol=1e-8, atol=1e-10
(the parity criterion of the benchmark), and the step sequence of the
oracle's explicit_rk restatement (attempt counts equal)."""
import numpy as np
import pytest
from numpy.testing import assert_allclose

pytestmark = pytest.mark.gpu

import qutip_amd as qa  # noqa: E402
from oracle import cbind, models, tableaus  # noqa: E402
from tests import systems  # noqa: E402

OPTS = {"rtol": 1e-8, "atol": 1e-10}
TOL = 1e-6


def _max_err(res, ref):
    return max(np.abs(np.asarray(res.expect[k]).real - ref[k]).max() for k in range(len(ref)))


@pytest.mark.parametrize("fused", [True, False])
def test_jc_mesolve_matches_golden(golden, fused):
    print(f"jc mesolve fused={fused}", flush=True)
    g = golden("jc_mesolve.npz")
    s = systems.jaynes_cummings()
    res = qa.mesolve(s["H"], s["psi0"], g["tlist"], s["c_ops"], e_ops=s["e_ops"],
                     options={**OPTS, "fused": fused})
    err = _max_err(res, g["expect"])
    print(f"jc fused={fused} max err {err:.3e}", flush=True)
    assert err < TOL
    assert np.asarray(res.expect[0]).dtype == np.float64  # real casting


def test_jc_step_sequence_matches_oracle():
    s = systems.jaynes_cummings()
    solver = qa.MESolver(s["H"], s["c_ops"], options=OPTS)
    solver.run(s["psi0"], s["tlist"], e_ops=s["e_ops"])
    st = solver._integrator.stats()
    m = models.jaynes_cummings()
    Lm = models.liouvillian(m["H"], m["c_ops"])
    _, _, ost = cbind.rk_solve(Lm, models.vec(models.ket2dm(m["psi0"])), m["tlist"],
                               tableaus.vern7, cbind.functionals_from_ops(m["e_ops"], 20),
                               rtol=1e-8, atol=1e-10)
    print("device", st, "oracle", ost, flush=True)
    assert st["attempts"] == ost["n_attempts"]
    assert st["rhs_evals"] == ost["n_rhs"]


def test_ising8_mesolve_matches_golden(golden):
    print("ising L=8 mesolve", flush=True)
    g = golden("ising8_mesolve.npz")
    s = systems.ising(8, 0.1)
    res = qa.mesolve(s["H"], s["psi0"], g["tlist"], s["c_ops"], e_ops=s["e_ops"], options=OPTS)
    err = _max_err(res, g["expect"])
    print(f"ising8 max err {err:.3e} stats {res.stats}", flush=True)
    assert err < TOL


def test_ising8_mesolve_fp32_operator(golden):
    g = golden("ising8_mesolve.npz")
    s = systems.ising(8, 0.1)
    res = qa.mesolve(s["H"], s["psi0"], g["tlist"], s["c_ops"], e_ops=s["e_ops"],
                     options={**OPTS, "operator_precision": "fp32"})
    err = _max_err(res, g["expect"])
    print(f"ising8 fp32-operator max err {err:.3e}", flush=True)
    # the <O> bar of the chosen precision form: 1e-6 against the FP64 fixture
    assert err < 1e-6


@pytest.mark.parametrize("tab,interp", [("tsit5", True), ("vern7", False)])
def test_jc_mesolve_other_forms(golden, tab, interp):
    g = golden("jc_mesolve.npz")
    s = systems.jaynes_cummings()
    res = qa.mesolve(s["H"], s["psi0"], g["tlist"], s["c_ops"], e_ops=s["e_ops"],
                     options={**OPTS, "tableau": tab, "interpolate": interp, "nsteps": 100000})
    assert _max_err(res, g["expect"]) < TOL


@pytest.mark.parametrize("fused", [True, False])
def test_ising6_sesolve_matches_golden(golden, fused):
    g = golden("ising6_sesolve.npz")
    s = systems.ising(6, 0.1)
    res = qa.sesolve(s["H"], s["psi0"], g["tlist"], e_ops=s["e_ops"],
                     options={**OPTS, "fused": fused})
    err = _max_err(res, g["expect"])
    print(f"ising6 sesolve fused={fused} err {err:.3e}", flush=True)
    assert err < TOL


def test_mesolve_decay_closed_form():
    """TestMESolveDecay (qutip/tests/solver/test_mesolve.py:84-95)."""
    N, kappa = 10, 0.2
    a = qa.destroy(N)
    tlist = np.linspace(0, 10, 101)
    res = qa.mesolve(a.dag() @ a, qa.basis(N, 9), tlist, [np.sqrt(kappa) * a],
                     e_ops=[a.dag() @ a], options=OPTS)
    assert_allclose(res.expect[0], 9 * np.exp(-kappa * tlist), atol=1e-6)


def test_too_much_work_raises():
    s = systems.jaynes_cummings()
    with pytest.raises(qa.IntegratorException, match="Too much work"):
        qa.mesolve(s["H"], s["psi0"], [0, 25], s["c_ops"], e_ops=s["e_ops"],
                   options={**OPTS, "nsteps": 2})


def test_final_state_and_states(golden):
    s = systems.jaynes_cummings()
    tl = np.linspace(0, 2, 5)
    res = qa.mesolve(s["H"], s["psi0"], tl, s["c_ops"], e_ops=s["e_ops"],
                     options={**OPTS, "store_final_state": True})
    rho = res.final_state.full()
    assert abs(np.trace(rho) - 1) < 1e-10
    assert_allclose(rho, rho.conj().T, atol=1e-10)
    res2 = qa.mesolve(s["H"], s["psi0"], tl, s["c_ops"], options={**OPTS, "store_states": True})
    assert len(res2.states) == len(tl)
    assert_allclose(res2.states[-1].full(), rho, atol=1e-9)


@pytest.mark.parametrize("fused", [True, False])
def test_integrator_protocol_qobjevo(fused):
    """TestIntegratorCallable, QobjEvo case (test_integrator.py:45-98)."""
    rng = np.random.default_rng(0)
    M = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
    mat = qa.Qobj(-1j * (M + M.conj().T))
    der = qa.QobjEvo(mat).matmul_data
    integ = qa.IntegratorHipRK(der, {"fused": fused})
    y0 = np.array([1, 1.0 + 0.5j, 0.5 - 0.5j])
    from scipy.linalg import expm
    analytical = lambda t: expm(t * mat.full()) @ y0
    tlist = np.linspace(0, 2, 21)
    integ.set_state(tlist[0], qa.HipDense.from_array(y0))
    for (t, st), t_in in zip(integ.run(tlist), tlist[1:]):
        assert t == t_in
        assert_allclose(st.to_array().ravel(), analytical(t), atol=5e-5)
    t, st = integ.integrate(3)
    assert t == 3.0
    assert_allclose(st.to_array().ravel(), analytical(3.0), atol=5e-5)
    t_front, _ = integ.mcstep(5)
    assert 3.0 < t_front <= 5.0
    t_target = (3 + t_front) / 2
    t_in, st = integ.mcstep(t_target)
    assert t_in == t_target
    assert_allclose(st.to_array().ravel(), analytical(t_in), atol=5e-5)


def test_profiling_leaves_the_engine_reusable():
    """The measurement entry points run stage kernels outside a solve; after
    a new set_state the engine must integrate exactly as a fresh one."""
    import ctypes as C

    from qutip_amd import _lib
    from qutip_amd.solver.hip_rk import RKEngine
    m = models.jaynes_cummings()
    Lm = models.liouvillian(m["H"], m["c_ops"])
    rho0 = models.vec(models.ket2dm(m["psi0"]))
    times = m["tlist"][1:20]

    def run(profile):
        eng = RKEngine(Lm, "vern7", 1e-8, 1e-10)
        eng.set_state(0.0, rho0)
        if profile:
            ms = np.zeros(16)
            _lib.call("qa_rk_profile_stages", eng.handle, 2,
                      ms.ctypes.data_as(C.POINTER(C.c_double)))
            t_step, t_dense = C.c_double(), C.c_double()
            _lib.call("qa_rk_profile_sequence", eng.handle, 3, C.byref(t_step), C.byref(t_dense))
            assert np.all(ms > 0) and t_step.value > 0 and t_dense.value > 0
            eng.set_state(0.0, rho0)
        eng.run(times, states_out=None)
        _, y = eng.get_state()
        return y.to_array()

    assert np.array_equal(run(False), run(True))
