"""mesolve / MESolver (qutip/solver/mesolve.py:24-277): Lindblad master
equation in the column-stacked superoperator form,
d vec(rho)/dt = L vec(rho) with L = liouvillian(H, c_ops)
(qutip/core/superoperator.py:51-142)."""
import numpy as np

from ..core import LindbladMatrixForm, Qobj, QobjEvo, liouvillian
from .hip_rk import IntegratorHipRK
from .sesolve import sesolve
from .solver_base import Solver

__all__ = ["mesolve", "MESolver"]


class MESolver(Solver):
    """With ``options={"matrix_form": True}`` the state stays an N x N matrix
    and the right-hand side is LindbladMatrixForm (mesolve.py:251-277): the
    dense path: MFMA products (``dense_precision`` 'fp32', 'bf16x6' or
    'bf16x3') or,
    for banded operators, FP64 products of their diagonals ('fp64_banded';
    'auto', the default, picks it when it applies)."""
    name = "mesolve"
    _avail_integrators = {}
    _preserves_hermiticity = True
    solver_options = {**Solver.solver_options, "matrix_form": False}

    def __init__(self, H, c_ops=None, *, options=None):
        if isinstance(H, (list, tuple)):
            H = QobjEvo(H)
        if isinstance(H, QobjEvo) and not H.elements:
            H = H.qobj
        if any(isinstance(c, QobjEvo) and c.elements for c in (c_ops or [])):
            raise NotImplementedError("time-dependent collapse operators are not supported")
        c_ops = [c.qobj if isinstance(c, QobjEvo) else c for c in (c_ops or [])]
        for c in c_ops:
            if not isinstance(c, Qobj):
                raise TypeError("c_ops must be Qobj")
        self._num_collapse = len(c_ops)
        self._matrix_form = bool((options or {}).get("matrix_form", False))
        if self._matrix_form:
            if isinstance(H, QobjEvo):
                raise NotImplementedError("matrix_form with a time-dependent H")
            if H.issuper or any(c.issuper for c in c_ops):
                raise ValueError("matrix_form=True cannot be used with superoperator H")
            self._vectorize_state = False
            rhs = LindbladMatrixForm(H, c_ops)
        else:
            rhs = liouvillian(H, [c for c in c_ops if not c.issuper])
            for c in c_ops:
                if c.issuper:
                    rhs = rhs + c
            rhs = rhs if isinstance(rhs, QobjEvo) else QobjEvo(rhs)
        super().__init__(rhs, options=options)


    def _prepare_state(self, state):
        # the matrix form's fast path assumes a Hermitian rho (A + A^dag);
        # an operator state that is not Hermitian switches it off
        # (mesolve.py:270-276).  A ket becomes ket2dm(psi), which is
        # Hermitian, so it switches the fast path back on (the reference
        # leaves the switch as the last operator state set it).
        if self._matrix_form and state.isoper:
            self.rhs.assume_hermitian_state = state.isherm
        elif self._matrix_form and state.isket:
            self.rhs.assume_hermitian_state = True
        return super()._prepare_state(state)


MESolver.add_integrator(IntegratorHipRK, "hip_rk")


def mesolve(H, rho0, tlist, c_ops=None, e_ops=None, args=None, options=None, **kwargs):
    """Master-equation evolution; a ket without collapse operators goes to
    sesolve like the reference (mesolve.py:155-162)."""
    c_ops = list(c_ops or [])
    if isinstance(H, (list, tuple)):
        H = QobjEvo(H, args=args, tlist=tlist)
    if not c_ops and isinstance(rho0, Qobj) and rho0.isket and not H.issuper:
        return sesolve(H, rho0, tlist, e_ops=e_ops, args=args, options=options)
    solver = MESolver(H, c_ops, options=options)
    return solver.run(rho0, tlist, e_ops=e_ops, args=args)
