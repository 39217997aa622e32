"""numpy restatement of QuTiP's stochastic Schroedinger / master equation
solvers (test oracle).

Follows:
* StochasticClosedSystem (ssesolve)   qutip/solver/sode/ssystem.pyx:119-176
    drift(psi)  = (-iH - 1/2 sum c^dag c) psi
                  + sum_i (e_i / 2 c_i psi - e_i^2 / 8 psi),
    e_i = <psi|c_i + c_i^dag|psi>,  diffusion_i(psi) = c_i psi - e_i / 2 psi
* StochasticOpenSystem (smesolve)     ssystem.pyx:195-250
    drift = liouvillian(H, sc_ops + c_ops) vec(rho),
    diffusion_i = C_i vec(rho) - tr(C_i rho) vec(rho), C_i = spre(c_i) +
    spost(c_i^dag)
* heterodyne detection                stochastic.py:269-274 (each sc_op as
                                      c / sqrt 2 and -i c / sqrt 2)
* the Euler and Platen steps           sode/_sode.pyx:17-120
* the steps per output interval        sode/sode.py:155-177 (N = floor of
                                      delta_t / dt, one more when the rest is
                                      above dt / 2; t advances by N dt)
* the Wiener increments               sode/_noise.py:8-30: normal(0, sqrt dt)
                                      of shape (N, 1, n_sc) from the
                                      trajectory's generator, in order.

Pinned by: the ensemble mean against the master equation (weak order of
the schemes) on a dephasing qubit and a damped cavity, and the closed form
of pure-dephasing measurement (<sigma_x> = e^{-2 gamma t} on average).
"""
import numpy as np
import scipy.sparse as sp

from . import models

CD = np.complex128


def step_counts(tlist, dt):
    """Steps of every output interval and the times the steps start from
    (sode.py:155-177)."""
    t = float(tlist[0])
    counts, starts = [], []
    for tk in tlist[1:]:
        delta = tk - t
        if delta < 0.5 * dt:
            counts.append(0)
            starts.append(t)
            continue
        N, extra = np.divmod(delta, dt)
        N = int(N)
        if extra > 0.5 * dt:
            N += 1
        counts.append(N)
        starts.append(t)
        t += dt * N
    return counts, starts


def wiener(seed, nsteps, nsc, dt):
    """The increments a trajectory consumes: generator.normal(0, sqrt(dt),
    (nsteps, 1, nsc)) of default_rng(seed) (_noise.py:15-30)."""
    gen = np.random.default_rng(seed)
    return gen.normal(0, np.sqrt(dt), size=(nsteps, 1, nsc))[:, 0, :]


class System:
    """The drift and diffusions of ssesolve (open=False, vectors psi) or
    smesolve (open=True, vec(rho))."""

    def __init__(self, H, sc_ops, c_ops=(), open=False, heterodyne=False):
        H = sp.csr_matrix(H, dtype=CD)
        sc = [sp.csr_matrix(c, dtype=CD) for c in sc_ops]
        if heterodyne:
            sc = [s for c in sc for s in (c / np.sqrt(2), c * (-1j / np.sqrt(2)))]
        self.open = open
        self.d = H.shape[0]
        if open:
            self.L = models.liouvillian(H, sc + [sp.csr_matrix(c, dtype=CD) for c in c_ops])
            eye = sp.identity(self.d, dtype=CD, format="csr")
            # spre(c) + spost(c^dag): vec(c rho + rho c^dag) (superoperator.py:379-415)
            self.M = [sp.csr_matrix(sp.kron(eye, c) + sp.kron(c.conj(), eye)) for c in sc]
        else:
            if c_ops:
                raise ValueError("c_ops are not supported by ssesolve")
            L = -1j * H
            for c in sc:
                L = L - 0.5 * (c.conj().T @ c)
            self.L = sp.csr_matrix(L)
            self.M = sc
        self.nsc = len(self.M)

    def _s(self, x, y):
        if self.open:
            return complex(np.sum(y[::self.d + 1]))  # trace_oper_ket
        return complex(np.vdot(x, y) + np.conj(np.vdot(x, y)))  # <x|c + c^dag|x>

    def drift(self, x):
        out = self.L @ x
        if not self.open:
            for M in self.M:
                y = M @ x
                e = self._s(x, y)
                out = out + (-0.125 * e * e) * x + (0.5 * e) * y
        return out

    def diffusion(self, x):
        out = []
        for M in self.M:
            y = M @ x
            e = self._s(x, y)
            out.append(y - (e if self.open else 0.5 * e) * x)
        return out


def euler_step(sys, x, dt, dW):
    """_sode.pyx:17-56."""
    a = sys.drift(x)
    b = sys.diffusion(x)
    new = x + dt * a
    for i in range(sys.nsc):
        new = new + dW[i] * b[i]
    return new


def platen_step(sys, x, dt, dW):
    """_sode.pyx:58-120 (Breuer & Petruccione eq. 7.47)."""
    sqrt_dt = np.sqrt(dt)
    sqrt_dt_inv = 0.25 / sqrt_dt
    d1 = x + dt * sys.drift(x)
    d2 = sys.diffusion(x)
    out = 0.5 * d1
    Vt = d1.copy()
    Vp, Vm = [], []
    for i in range(sys.nsc):
        Vp.append(d1 + sqrt_dt * d2[i])
        Vm.append(d1 - sqrt_dt * d2[i])
        Vt = Vt + dW[i] * d2[i]
    out = out + 0.5 * dt * sys.drift(Vt)
    out = out + 0.5 * x
    for i in range(sys.nsc):
        d2p = sys.diffusion(Vp[i])
        d2m = sys.diffusion(Vm[i])
        dw = dW[i] * 0.25
        out = out + 2 * dw * d2[i]
        for j in range(sys.nsc):
            if i == j:
                dw2 = sqrt_dt_inv * (dW[i] * dW[j] - dt)
                dw2p, dw2m = dw2 + dw, -dw2 + dw
            else:
                dw2p = sqrt_dt_inv * dW[i] * dW[j]
                dw2m = -dw2p
            out = out + dw2p * d2p[j] + dw2m * d2m[j]
    return out


def trajectory(sys, x0, tlist, dt, dW, e_ops, method="platen"):
    """<O>(t) of one trajectory at every time of tlist: <x|O|x> for
    ssesolve, tr(O rho) for smesolve (not normalised, the stochastic
    solvers' normalize_output default)."""
    step = platen_step if method == "platen" else euler_step
    counts, _ = step_counts(tlist, dt)
    x = np.asarray(x0, dtype=CD).copy()
    d = sys.d

    def ev(x):
        if sys.open:
            rho = x.reshape((d, d), order="F")
            return [complex(np.sum(sp.csr_matrix(O).multiply(rho.T))) for O in e_ops]
        return [complex(np.vdot(x, O @ x)) for O in e_ops]
    out = [ev(x)]
    k = 0
    for N in counts:
        for _ in range(N):
            x = step(sys, x, dt, dW[k])
            k += 1
        out.append(ev(x))
    return np.array(out), x
