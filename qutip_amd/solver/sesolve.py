"""sesolve / SESolver (qutip/solver/sesolve.py:21-163): d psi / dt = -i H psi."""
import numpy as np

from ..core import Qobj, QobjEvo
from .hip_rk import IntegratorHipRK
from .solver_base import Solver

__all__ = ["sesolve", "SESolver"]


class SESolver(Solver):
    name = "sesolve"
    _avail_integrators = {}
    _preserves_hermiticity = False

    def __init__(self, H, *, options=None):
        if isinstance(H, (list, tuple)):
            H = QobjEvo(H)
        if not isinstance(H, (Qobj, QobjEvo)):
            raise TypeError("The Hamiltonian must be a Qobj or QobjEvo")
        rhs = H if H.issuper else -1j * H  # rhs = -iH (sesolve.py:160)
        super().__init__(QobjEvo(rhs), options=options)


SESolver.add_integrator(IntegratorHipRK, "hip_rk")


def sesolve(H, psi0, tlist, e_ops=None, args=None, options=None, **kwargs):
    """Schroedinger equation evolution of a ket or a unitary; H may be a
    time-dependent QobjEvo or its list form (sesolve.py:21-100)."""
    if isinstance(H, (list, tuple)):
        H = QobjEvo(H, args=args, tlist=tlist)
    solver = SESolver(H, options=options)
    return solver.run(psi0, tlist, e_ops=e_ops, args=args)
