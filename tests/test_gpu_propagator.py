"""propagator / Propagator on the GPU (qutip/solver/propagator.py:110-420,
mirroring qutip/tests/solver/test_propagator.py): U(t) is the solution from
the identity, run by the fused engine as one block state (kron(I_m, A) over
the column-major buffer).

* sesolve propagator of a random Hermitian H against expm(-iHt) (and
  unitary), the same step-attempt count as the oracle's RK on the
  flattened matrix;
* the superoperator propagator of a damped oscillator against expm(L t);
* Propagator: cached times, U(t, t_start) = U(t - t_start) for constant
  systems, inv;
* a time-dependent H (string coefficient) against the oracle's RK on the
  same list-form right-hand side."""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

import qutip_amd as qa  # noqa: E402
from oracle import models, rk, tableaus  # noqa: E402

OPTS = {"rtol": 1e-8, "atol": 1e-10}


def _rand_h(d, seed):
    rng = np.random.default_rng(seed)
    h = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    return (h + h.conj().T) / 4


def test_sesolve_propagator_vs_expm():
    d = 20
    H = _rand_h(d, 1)
    tl = np.linspace(0, 2, 5)
    print(f"sesolve propagator d={d}", flush=True)
    Us = qa.propagator(qa.Qobj(H), tl, options=OPTS)
    assert len(Us) == len(tl)
    err = max(np.abs(U.full() - sla.expm(-1j * H * t)).max() for U, t in zip(Us, tl))
    uni = max(np.abs(U.full() @ U.full().conj().T - np.eye(d)).max() for U in Us)
    print(f"  max |U - expm| {err:.2e}, unitarity {uni:.2e}", flush=True)
    assert err < 1e-6 and uni < 1e-6
    U1 = qa.propagator(qa.Qobj(H), 2.0, options=OPTS)
    assert np.abs(U1.full() - Us[-1].full()).max() < 1e-12


def test_propagator_runs_the_block_on_the_fused_engine():
    """One block state: the engine's attempts equal the oracle's RK on
    vec(U) with dU/dt = -iH U (the reference integrates the Dense matrix)."""
    d = 12
    H = _rand_h(d, 2)
    solver = qa.SESolver(qa.Qobj(H), options={**OPTS, "store_states": True})
    res = solver.run(qa.Qobj(np.eye(d)), [0.0, 1.5])
    integ = solver._integrator
    assert integ._engine is not None and integ._engine.n == d * d
    st = integ._engine.stats()
    A = sp.kron(sp.identity(d), sp.csr_matrix(-1j * H), format="csr")
    _, ork = rk.solve(lambda t, y: A @ y, np.eye(d).ravel(order="F"), [0.0, 1.5],
                      tableaus.TABLEAUS["vern7"], **OPTS)
    print(f"block attempts {st['attempts']} vs oracle {ork.n_attempts}", flush=True)
    assert st["attempts"] == ork.n_attempts
    assert np.abs(res.states[-1].full() - sla.expm(-1.5j * H)).max() < 1e-6


def test_mesolve_propagator_vs_expm():
    N = 4
    a = qa.destroy(N)
    H = 0.5 * a.dag() @ a + 0.3 * (a + a.dag())
    c_ops = [np.sqrt(0.2) * a]
    print("mesolve propagator", flush=True)
    U = qa.propagator(H, 1.3, c_ops, options=OPTS)
    L = models.liouvillian(H.to_csr(), [c.to_csr() for c in c_ops]).toarray()
    want = sla.expm(L * 1.3)
    err = np.abs(U.full() - want).max()
    print(f"  max |U - expm(L t)| {err:.2e}", flush=True)
    assert U.issuper and err < 1e-6


def test_propagator_class():
    d = 8
    H = qa.Qobj(_rand_h(d, 3))
    P = qa.Propagator(H, options=OPTS)
    U1 = P(1.0)
    assert P(1.0 + 1e-15) is U1  # within tol: the kept one
    U2 = P(2.5)
    assert np.abs(U2.full() - sla.expm(-2.5j * H.full())).max() < 1e-6
    U21 = P(2.5, 1.0)  # constant: U(1.5)
    assert np.abs(U21.full() - sla.expm(-1.5j * H.full())).max() < 1e-6
    assert np.abs((P.inv(1.0) @ U1).full() - np.eye(d)).max() < 1e-6


def test_time_dependent_propagator_vs_oracle():
    sx, sz = qa.sigmax(), qa.sigmaz()
    H = [0.7 * sz, [sx, "cos(w * t)"]]
    tl = [0.0, 1.0, 2.0]
    print("time-dependent propagator", flush=True)
    Us = qa.propagator(H, tl, args={"w": 2.0}, options=OPTS)
    f = models.str_coefficient("cos(w * t)", {"w": 2.0})
    A0 = sp.kron(sp.identity(2), -1j * 0.7 * sz.to_csr(), format="csr")
    A1 = sp.kron(sp.identity(2), -1j * sx.to_csr(), format="csr")
    rhs = models.td_rhs(A0, [(A1, f)])
    ys, _ = rk.solve(rhs, np.eye(2).ravel(order="F").astype(complex), tl,
                     tableaus.TABLEAUS["vern7"], rtol=1e-10, atol=1e-12)
    err = max(np.abs(U.full() - y.reshape(2, 2, order="F")).max() for U, y in zip(Us, ys))
    print(f"  max |U - oracle| {err:.2e}", flush=True)
    assert err < 1e-6
