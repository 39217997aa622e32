"""Solvers of the hot path: sesolve / mesolve (fused or plain hip_rk) and
the batched mcsolve."""
from .integrator import Integrator, IntegratorException
from .hip_rk import IntegratorHipRK, RKEngine
from .result import Result
from .solver_base import Solver
from .sesolve import sesolve, SESolver
from .mesolve import mesolve, MESolver
from .mcsolve import mcsolve, MCSolver, McResult
from .propagator import propagator, Propagator
from .steadystate import steadystate
from .stochastic import ssesolve, smesolve, SSESolver, SMESolver, StochasticResult

Solver.add_integrator(IntegratorHipRK, "hip_rk")
