"""Step-only iterations of the fused engine against full iterations.

While the device loop steps towards an output time that is several attempts
away, the host enqueues iterations that end after rk_control_kernel (no
dense-output stages, no output launch); an attempt that reaches the output
leaves the loop in the dense-output phase, the step-only iterations after it
return at once, and the next full iteration carries on.  With
QUTIP_AMD_STEP_ONLY=0 (read when an engine is created) every iteration is
full.  Both must give the same bits: the same attempts, the same dense-output
preparations, the same outputs and final state, on graph replays and on
direct launches, with and without stored states.  Bar: bit-identical."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import qutip_amd as qa  # noqa: E402
from qutip_amd.solver.hip_rk import RKEngine  # noqa: E402
from oracle import cbind, models  # noqa: E402
from tests import systems  # noqa: E402

STATS = ("attempts", "accepted", "dense_preps", "rhs_evals", "status", "out_idx")


def _pair(monkeypatch, make):
    """(full iterations only, step-only iterations allowed)."""
    monkeypatch.setenv("QUTIP_AMD_STEP_ONLY", "0")
    full = make()
    monkeypatch.delenv("QUTIP_AMD_STEP_ONLY")
    return full, make()


def _solve(eng, y0, tl, fun, store):
    fptr, fidx, fw = fun
    eng.set_functionals(np.zeros(len(fptr) - 1, np.int32), fptr.astype(np.int32),
                        np.zeros(len(fidx), np.int32), fidx.astype(np.int32),
                        fw.view(np.complex128))
    eng.set_state(tl[0], y0)
    buf = qa.HipDense((eng.npad, len(tl) - 1)) if store else None
    vals, st, _ = eng.run(tl[1:], states_out=buf.ptr if store else None)
    assert st >= 0
    t, y = eng.get_state()
    stats = eng.stats()
    return (vals, t, y.to_array().ravel(), buf.to_array() if store else None,
            {k: stats[k] for k in STATS}, stats["graph_launches"])


def _same(a, b):
    assert a[4] == b[4]
    assert np.array_equal(a[0], b[0])
    assert a[1] == b[1]
    assert np.array_equal(a[2], b[2])
    if a[3] is not None:
        assert np.array_equal(a[3], b[3])


TLISTS = {
    "sparse outputs": np.linspace(0.0, 20.0, 5),
    "dense then sparse": np.concatenate([np.linspace(0.0, 0.5, 11), [4.0, 20.0]]),
}


@pytest.mark.parametrize("graph", [True, False])
@pytest.mark.parametrize("store", [False, True])
def test_step_only_iterations_sparse_engine(monkeypatch, graph, store):
    print(f"step-only iterations: TFIM L=4 Liouvillian, graph {graph}, store {store}", flush=True)
    if not graph:
        monkeypatch.setenv("QUTIP_AMD_GRAPH", "0")
    m = models.ising(4, 0.1)
    Lm = models.liouvillian(m["H"], m["c_ops"]).tocsr()
    d = m["H"].shape[0]
    rho0 = models.vec(models.ket2dm(m["psi0"]))
    fun = cbind.functionals_from_ops(m["e_ops"], d)
    full, fast = _pair(monkeypatch, lambda: RKEngine(Lm, "vern7", 1e-8, 1e-10, lanes=1))
    for name, tl in TLISTS.items():
        a = _solve(full, rho0, tl, fun, store)
        b = _solve(fast, rho0, tl, fun, store)
        print(f"  {name}: {a[4]['attempts']} attempts, iterations launched {a[5]} (full) "
              f"{b[5]} (step-only allowed)", flush=True)
        _same(a, b)
        # again on the same engines: the plan restarts with every run
        _same(a, _solve(fast, rho0, tl, fun, store))


@pytest.mark.parametrize("prec", ["fp64_banded", "fp32"])
def test_step_only_iterations_matrix_form(monkeypatch, prec):
    print(f"step-only iterations: Kerr N=32 matrix form, {prec}", flush=True)
    s = systems.kerr(32)
    tl = np.linspace(0.0, 6.0, 4)
    opts = {"matrix_form": True, "dense_precision": prec, "rtol": 1e-8, "atol": 1e-10,
            "store_final_state": True}

    def solve():
        solver = qa.MESolver(s["H"], s["c_ops"], options=opts)
        res = solver.run(s["psi0"], tl, e_ops=s["e_ops"])
        st = solver._integrator.stats()
        return np.array(res.expect), res.final_state.full(), {k: st[k] for k in STATS}

    monkeypatch.setenv("QUTIP_AMD_STEP_ONLY", "0")
    a = solve()
    monkeypatch.delenv("QUTIP_AMD_STEP_ONLY")
    b = solve()
    print(f"  {a[2]['attempts']} attempts", flush=True)
    assert a[2] == b[2]
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1], b[1])
