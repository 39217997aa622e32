"""propagator / Propagator (qutip/solver/propagator.py:110-420): the
evolution operator U(t) with psi(t) = U(t) psi(0), or the superoperator
with vec(rho(t)) = U(t) vec(rho(0)).

As in the reference, U(t) is the solution of the equation of motion from
the identity: sesolve of -iH on the d x d identity, mesolve of the
Liouvillian on the d^2 x d^2 identity.  The hip_rk integrator runs such a
block of columns as one state on its fused engine (the block-diagonal
generator kron(I_m, A) over the column-major buffer), so the whole matrix
takes one step-size sequence and one error norm, as the reference's Dense
state does."""
import numbers

import numpy as np

from ..core import Qobj, QobjEvo, liouvillian, qeye_like
from .mesolve import MESolver, mesolve
from .sesolve import SESolver, sesolve

__all__ = ["propagator", "Propagator"]


def propagator(H, t, c_ops=None, args=None, options=None, **kwargs):
    """U(t) from 0 to t (a number), or from t[0] to every time of the list t
    (then the first output is the identity) (propagator.py:110-195)."""
    if isinstance(t, numbers.Real):
        tlist, list_output = [0.0, float(t)], False
    else:
        tlist, list_output = list(t), True
    if not isinstance(H, (Qobj, QobjEvo)):
        H = QobjEvo(H, args=args, **kwargs)
    if c_ops is not None and not isinstance(c_ops, (list, tuple)):
        c_ops = [c_ops]
    c_ops = [c if isinstance(c, (Qobj, QobjEvo)) else QobjEvo(c, args=args, **kwargs)
             for c in (c_ops or [])]
    if c_ops:
        H = liouvillian(H, c_ops)
    U0 = qeye_like(H.qobj if isinstance(H, QobjEvo) else H)
    if H.issuper:
        out = mesolve(H, U0, tlist, args=args, options=options).states
    else:
        out = sesolve(H, U0, tlist, args=args, options=options).states
    return out if list_output else out[-1]


class Propagator:
    """A propagator generator: ``U = Propagator(H, c_ops=...)``, then
    ``U(t)`` or ``U(t, t_start)`` (propagator.py:230-420).  Computed
    propagators are kept, sorted by time (up to ``memoize``), and a new one
    is evolved from the nearest earlier one; times within ``tol`` of a kept
    one reuse it.  For a constant system U(t, t_start) = U(t - t_start),
    else U(t) U(t_start)^-1."""

    def __init__(self, system, *, c_ops=None, args=None, options=None, memoize=10, tol=1e-14):
        if isinstance(system, (SESolver, MESolver)):
            self.solver = system
        else:
            Hevo = QobjEvo(system, args=args) if not isinstance(system, QobjEvo) else system
            c_ops = [c if isinstance(c, (Qobj, QobjEvo)) else QobjEvo(c, args=args)
                     for c in (c_ops or [])]
            if Hevo.issuper or c_ops:
                self.solver = MESolver(Hevo, c_ops=c_ops, options=options)
            else:
                self.solver = SESolver(Hevo, options=options)
        self.solver.options = {**self.solver.options, "store_states": True}
        rhs = self.solver.rhs
        self.cte = getattr(rhs, "isconstant", True)
        H0 = rhs.qobj if isinstance(rhs, QobjEvo) else rhs
        self.unitary = isinstance(self.solver, SESolver) and (1j * H0).isherm
        self.times = [0.0]
        self.props = [qeye_like(H0)]
        self.args = args
        self.memoize = max(3, int(memoize))
        self.tol = tol

    def _inv(self, U):
        if self.unitary:
            return U.dag()
        return Qobj(np.linalg.inv(U.full()), dims=U.dims)

    def _lookup_or_compute(self, t):
        idx = int(np.searchsorted(self.times, t))
        if idx < len(self.times) and abs(t - self.times[idx]) <= self.tol:
            return self.props[idx]
        if idx > 0 and abs(t - self.times[idx - 1]) <= self.tol:
            return self.props[idx - 1]
        if idx == 0:
            raise ValueError("propagators are computed forward from t = 0")
        res = self.solver.run(self.props[idx - 1], [self.times[idx - 1], t], args=None)
        U = res.states[-1]
        if len(self.times) < self.memoize:
            self.times.insert(idx, t)
            self.props.insert(idx, U)
        return U

    def __call__(self, t, t_start=0.0, **args):
        if not self.cte and args and args != self.args:
            self.args = args
            self.solver._argument(args)
            self.times, self.props = [0.0], [self.props[0]]
        if t_start:
            if self.cte:
                return self._lookup_or_compute(t - t_start)
            return self._lookup_or_compute(t) @ self._inv(self._lookup_or_compute(t_start))
        return self._lookup_or_compute(t)

    def inv(self, t, **args):
        """U(t)^-1 (propagator.py:400-420)."""
        return self._inv(self(t, **args))
