"""FP64 reference solutions for constant generators (test oracle).

Test infrastructure only: nothing in the product imports this module.

For a time-independent right-hand side dy/dt = A y the exact solution is
y(t) = expm(A (t - t0)) y(t0); scipy.sparse.linalg.expm_multiply (Al-Mohy &
Higham 2011) evaluates it to ~1e-12 on a uniform grid.  This is the library
routine the parity tests compare the solvers against.

The output step follows qutip/solver/solver_base.py:138-157 (normalise by
tr(rho) or ||psi|| when the input was normalised) and qutip/core/expect.py:79-101
(tr(O rho) / <psi|O|psi>, real part for Hermitian O).
"""
import numpy as np
import scipy.sparse.linalg as spla

from . import models, rk, tableaus


def states_uniform(A, y0, tlist):
    """y(t) for every t of a uniform ``tlist`` (expm_multiply needs a grid)."""
    t0, tf, n = float(tlist[0]), float(tlist[-1]), len(tlist)
    if not np.allclose(np.diff(tlist), (tf - t0) / (n - 1), rtol=1e-12, atol=1e-14):
        raise ValueError("expm_multiply path needs a uniform tlist")
    ys = spla.expm_multiply(A * 1.0, np.asarray(y0, dtype=np.complex128),
                            start=0.0, stop=tf - t0, num=n, endpoint=True)
    return np.asarray(ys)


def mesolve_expect(H, c_ops, psi0, tlist, e_ops, normalize=True):
    """<O>(t) of the master equation with the restated Liouvillian."""
    d = H.shape[0]
    Lm = models.liouvillian(H, c_ops)
    rho0 = models.ket2dm(psi0) if psi0.ndim == 1 else psi0
    vs = states_uniform(Lm, models.vec(rho0), tlist)
    out = np.zeros((len(e_ops), len(tlist)))
    for k, v in enumerate(vs):
        rho = models.unvec(v, d)
        tr = np.trace(rho)
        for j, O in enumerate(e_ops):
            val = models.expect_dm(O, rho)
            if normalize:
                val = val / tr
            out[j, k] = val.real
    return out


def sesolve_expect(H, psi0, tlist, e_ops, normalize=True):
    vs = states_uniform(-1j * H, psi0, tlist)
    out = np.zeros((len(e_ops), len(tlist)))
    for k, psi in enumerate(vs):
        nrm2 = np.vdot(psi, psi).real
        for j, O in enumerate(e_ops):
            val = models.expect_ket(O, psi)
            if normalize:
                val = val / nrm2
            out[j, k] = val.real
    return out


def mesolve_expect_rk(H, c_ops, psi0, tlist, e_ops, rtol=1e-10, atol=1e-12, normalize=True):
    """<O>(t) of the master equation for systems whose superoperator is too
    large for expm_multiply (dense collapse operators make it d^4 dense):
    the restated vern7 integrator (oracle/rk.py) on the matrix-form
    right-hand side (models.matrix_form_rhs) at tight tolerances, FP64."""
    d = H.shape[0]
    rho0 = models.ket2dm(psi0) if np.ndim(psi0) == 1 else psi0
    f = models.matrix_form_rhs(H, c_ops)
    ys, _ = rk.solve(f, models.vec(rho0), np.asarray(tlist, float), tableaus.TABLEAUS["vern7"],
                     rtol=rtol, atol=atol, nsteps=10 ** 7)
    out = np.zeros((len(e_ops), len(tlist)))
    for k, v in enumerate(ys):
        rho = models.unvec(v, d)
        tr = np.trace(rho)
        for j, O in enumerate(e_ops):
            val = models.expect_dm(O, rho)
            out[j, k] = (val / tr if normalize else val).real
    return out
