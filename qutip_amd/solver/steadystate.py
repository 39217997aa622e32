"""steadystate (qutip/solver/steadystate.py:36-216): rho_ss with
L(rho_ss) = 0 and tr rho_ss = 1.

* method "direct" (the reference's default, _steadystate_direct :218-274):
  the first row of L gets w vec(I)^T added (w the mean absolute value of
  L's non-zeros), the right-hand side is w e_0, and the sparse system is
  solved by LU (scipy's spsolve, which the reference calls too); the
  solution is unstacked and made Hermitian.
* method "propagator" (_steadystate_expm :298-318): rho is propagated by
  U(T), U(2T), U(4T), ... (the propagator squared each round) until the
  Hilbert-Schmidt distance of successive iterates is below
  ``propagator_tol``.  Here every U(2^k T) rho is a fused-engine mesolve
  on the GPU (the whole step loop on the device) instead of a dense
  matrix exponential."""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from ..core import Qobj, QobjEvo, liouvillian, rand_dm
from .mesolve import MESolver

__all__ = ["steadystate"]


def _as_liouvillian(A, c_ops):
    if isinstance(A, QobjEvo):
        if A.elements:
            raise ValueError("steadystate needs a constant generator")
        A = A.qobj
    if not A.issuper or c_ops:
        A = liouvillian(A, list(c_ops or []))
    return A


def _oper_dims(L):
    return L.dims[0]


def _steadystate_direct(L, weight=0.0):
    A = sp.csr_matrix(L.to_csr(), dtype=np.complex128)
    N = A.shape[0]
    n = int(round(N ** 0.5))
    if not weight:
        weight = float(np.mean(np.abs(A.data[A.data != 0])))
    vec_eye = sp.csr_matrix(np.eye(n, dtype=np.complex128).ravel(order="F")[None, :])
    A = sp.lil_matrix(A)
    A[0, :] = A[0, :] + weight * vec_eye
    b = np.zeros(N, dtype=np.complex128)
    b[0] = weight
    x = spla.spsolve(sp.csc_matrix(A), b)
    rho = x.reshape((n, n), order="F")
    rho = 0.5 * (rho + rho.conj().T)
    return Qobj(rho, dims=_oper_dims(L), isherm=True)


def _hilbert_dist(a, b):
    d = a - b
    return float(np.real(np.trace(d.conj().T @ d)))


def _steadystate_propagator(L, rho=None, propagator_tol=1e-5, propagator_T=10,
                            propagator_max_iter=30, options=None):
    d = int(round(L.shape[0] ** 0.5))
    dims = _oper_dims(L)
    if rho is None:
        rho = rand_dm(d)
    rho = Qobj(rho.full() if isinstance(rho, Qobj) else rho, dims=dims)
    solver = MESolver(L, [], options={"rtol": 1e-8, "atol": 1e-10, **(options or {}),
                                      "store_final_state": True})
    T = float(propagator_T)
    for _ in range(int(propagator_max_iter)):
        nxt = solver.run(rho, [0.0, T]).final_state.full()
        nxt = (nxt + nxt.conj().T) / (2 * np.trace(nxt))
        if _hilbert_dist(nxt, rho.full()) <= propagator_tol:
            return Qobj(nxt, dims=dims, isherm=True)
        rho = Qobj(nxt, dims=dims)
        T *= 2
    raise RuntimeError(f"Did not converge to a steadystate after {propagator_max_iter} "
                       "iterations.")


def steadystate(A, c_ops=(), *, method="direct", solver=None, **kwargs):
    """The steady state of a Hamiltonian with collapse operators, or of a
    Liouvillian (steadystate.py:36-216); methods "direct" and
    "propagator"."""
    L = _as_liouvillian(A, c_ops)
    if method in ("direct", "iterative"):
        if solver not in (None, "spsolve"):
            raise ValueError(f"solver {solver!r} is not supported (spsolve)")
        return _steadystate_direct(L, kwargs.pop("weight", 0.0))
    if method == "propagator":
        return _steadystate_propagator(L, **kwargs)
    raise ValueError(f"method {method} not supported.")
