"""qutip_amd -- MI355X-native (gfx950 HIP) time evolution for QuTiP solvers.

The hot path of sesolve / mesolve / mcsolve (the right-hand-side product and
the Runge-Kutta stage updates) runs in hand-written HIP kernels behind the C
ABI of include/qutip_amd.h.  Importing the package does not touch the GPU;
the first device operation loads libqutip_amd.so and fails loudly when it or
a GPU is missing.
"""
from . import _lib
from .core import (Qobj, QobjEvo, LindbladMatrixForm, basis, ket2dm, destroy, create, num, qeye, sigmax, sigmay,
                   sigmaz, sigmam, sigmap, tensor, spre, spost, liouvillian,
                   lindblad_dissipator, expect, rand_herm, rand_dm)
from .data import HipDense, HipCSR
from .fileio import file_data_store, file_data_read, qsave, qload
from .solver import (sesolve, mesolve, mcsolve, SESolver, MESolver, MCSolver, McResult, Solver,
                     Result, Integrator, IntegratorException, IntegratorHipRK, propagator,
                     Propagator, steadystate, ssesolve, smesolve, SSESolver, SMESolver,
                     StochasticResult)
from . import coefficient  # the module: coefficient.coefficient(...) builds one

__version__ = "0.1.0"
