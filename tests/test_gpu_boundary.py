"""The drop-in boundary's device paths.

* Solver.run's generic path (stored states, a callable e_op) drives
  Integrator.run(tlist), which on the fused engine is one device loop for
  the whole list (integrator.py:197-212 calls integrate per time): one
  engine run per Solver.run, states equal to the step-wise path;
* the whole-solve kernel at n = 512 with an ELL width of 8 and functionals
  (the launch whose LDS plan once exceeded 160 KB with its static tableau)
  against the matrix exponential."""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

import qutip_amd as qa  # noqa: E402
from qutip_amd.solver import hip_rk  # noqa: E402
from oracle import exact  # noqa: E402
from tests import systems  # noqa: E402

OPTS = {"rtol": 1e-8, "atol": 1e-10}


def test_generic_path_is_one_device_loop(monkeypatch):
    s = systems.ising(6, 0.1, tf=2.0, npts=9)  # Liouvillian 4096 x 4096: the fused engine
    calls = []
    orig = hip_rk.RKEngine.run

    def counted(self, *a, **k):
        calls.append(len(np.atleast_1d(a[0])))
        return orig(self, *a, **k)

    monkeypatch.setattr(hip_rk.RKEngine, "run", counted)
    solver = qa.MESolver(s["H"], s["c_ops"], options={**OPTS, "store_states": True})
    mz = s["e_ops"][0]
    res = solver.run(s["psi0"], s["tlist"], e_ops=[lambda t, st: qa.expect(mz, st)])
    print(f"engine runs {calls}", flush=True)
    assert calls == [len(s["tlist"]) - 1]
    # the same states step by step
    integ = solver._integrator
    integ.set_state(s["tlist"][0], solver._prepare_state(s["psi0"]))
    for t, st in zip(s["tlist"][1:], res.states[1:]):
        _, y = integ.integrate(t)
        assert np.abs(solver._restore_state(y).full() - st.full()).max() < 1e-12
    m = exact.mesolve_expect(s["H"].to_csr(), [c.to_csr() for c in s["c_ops"]],
                             s["psi0"].full().ravel(), s["tlist"], [mz.to_csr()])
    assert np.abs(np.asarray(res.expect[0]) - m[0]).max() < 1e-6


def test_whole_solve_n512_width8():
    d = 512
    rng = np.random.default_rng(4)
    offs = [0, 1, 2, 5]  # 7 diagonals with the mirrors: ELL width 8
    H = sp.csr_matrix((d, d), dtype=complex)
    for o in offs:
        v = rng.standard_normal(d - o) + (1j * rng.standard_normal(d - o) if o else 0)
        H = H + sp.diags(v, o) + (sp.diags(v.conj(), -o) if o else 0)
    H = qa.Qobj(H)
    w = int(np.diff(H.to_csr().indptr).max())
    assert w == 7
    print(f"whole solve n={d}, longest row {w}", flush=True)
    psi0 = qa.basis(d, 0)
    tl = np.linspace(0, 1, 11)
    e_ops = [qa.Qobj(sp.diags(np.arange(d) * 1.0)), H]
    solver = qa.SESolver(H, options=OPTS)
    res = solver.run(psi0, tl, e_ops=e_ops)
    assert solver._integrator._small is not None
    ref = exact.sesolve_expect(H.to_csr(), psi0.full().ravel(), tl, [O.to_csr() for O in e_ops])
    err = max(np.abs(np.asarray(res.expect[k]) - ref[k]).max() for k in range(2))
    print(f"  max err {err:.2e}", flush=True)
    assert err < 1e-6
