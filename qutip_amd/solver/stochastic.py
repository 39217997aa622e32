"""ssesolve / smesolve and their solvers (qutip/solver/stochastic.py:311-900)
with all trajectories in one GPU launch (qa_sde_run).

The systems are those of qutip/solver/sode/ssystem.pyx (StochasticClosed-
System :119-176, StochasticOpenSystem :195-250), the steps Euler and
Platen (sode/_sode.pyx:17-120; "platen" is the reference's default), the
number of steps per output interval that of _Explicit_Simple_Integrator.
integrate (sode/sode.py:155-177), and the Wiener increments those the
reference draws: normal(0, sqrt(dt), (N, 1, n_sc)) from each trajectory's
generator, in order (sode/_noise.py:8-30), with the seeds of
MultiTrajSolver (multitraj.py:354-393).  Heterodyne detection splits every
sc_op c into c / sqrt 2 and -i c / sqrt 2 (stochastic.py:269-274);
measurements are <m_op>(t) + dW_factor dW / dt with m_ops = c + c^dag
(and -i (c - c^dag) for heterodyne) (stochastic.py:686-705, 124-180).
"""
import ctypes as C
import time

import numpy as np
import scipy.sparse as sp
from numpy.random import SeedSequence

from .. import _lib
from .. import data as _data
from ..core import Qobj, QobjEvo, liouvillian, ket2dm

__all__ = ["ssesolve", "smesolve", "SSESolver", "SMESolver", "StochasticResult"]

METHODS = {"euler": 0, "platen": 1}
MAX_SC = 4
MAX_N = 1024
MAX_E = 16


def _qobj(op):
    if isinstance(op, QobjEvo):
        if op.elements:
            raise NotImplementedError("time-dependent stochastic systems are not supported")
        return op.qobj
    return op


def _ell(A, n, width=None):
    """Column-major ELL (entry u of row r at [u, r]) of width max(4, the
    longest row rounded up to 4) or ``width``; padding entries value 0 at
    their own row's column."""
    A = sp.csr_matrix(A, dtype=np.complex128)
    A.sum_duplicates()
    A.sort_indices()
    lens = np.diff(A.indptr)
    W = max(4, -(-int(lens.max() if n else 0) // 4) * 4)
    if width is not None:
        W = width
    col = np.tile(np.arange(n, dtype=np.int32), (W, 1))
    val = np.zeros((W, n), dtype=np.complex128)
    rows = np.repeat(np.arange(n), lens)
    slot = np.arange(A.nnz) - np.repeat(A.indptr[:-1], lens)
    col[slot, rows] = A.indices
    val[slot, rows] = A.data
    return W, col, val


def _csr_block(ops, rows, ncols):
    rowptr, cols, vals = [], [], []
    off = 0
    for O in ops:
        A = sp.csr_matrix(O, dtype=np.complex128)
        A.sort_indices()
        if A.shape != (rows, ncols):
            raise ValueError(f"e_op of shape {A.shape}, expected {(rows, ncols)}")
        rowptr.append(A.indptr.astype(np.int64) + off)
        cols.append(A.indices)
        vals.append(A.data)
        off += A.nnz
    rp = np.concatenate(rowptr).astype(np.int32) if rowptr else np.zeros(1, np.int32)
    cl = np.concatenate(cols).astype(np.int32) if off else np.zeros(1, np.int32)
    vl = np.concatenate(vals) if off else np.zeros(1, np.complex128)
    return rp, cl, vl


def step_counts(tlist, dt):
    """Steps of every output interval as _Explicit_Simple_Integrator.
    integrate takes them (sode/sode.py:155-177): N = floor(delta / dt), one
    more when the rest exceeds dt / 2, none below dt / 2; t advances by
    N dt."""
    t = float(tlist[0])
    counts = []
    for tk in tlist[1:]:
        delta = tk - t
        if delta < 0:
            raise ValueError("Integration time, can't be negative.")
        if delta < 0.5 * dt:
            counts.append(0)
            continue
        N, extra = np.divmod(delta, dt)
        N = int(N)
        if extra > 0.5 * dt:
            N += 1
        counts.append(N)
        t += dt * N
    return counts


def _read_seeds(seeds, ntraj):
    """MultiTrajSolver._read_seed (multitraj.py:354-374)."""
    if seeds is None:
        return SeedSequence().spawn(ntraj)
    if isinstance(seeds, SeedSequence):
        return seeds.spawn(ntraj)
    if not isinstance(seeds, list):
        return SeedSequence(seeds).spawn(ntraj)
    if len(seeds) >= ntraj:
        return [s if isinstance(s, SeedSequence) else SeedSequence(s) for s in seeds[:ntraj]]
    raise ValueError("A seed list must be longer than ntraj")


class StochasticResult:
    """Ensemble of stochastic trajectories (StochasticResult,
    stochastic.py:182-245, over MultiTrajResult): average and standard
    deviation of every e_op, optionally the runs, the Wiener increments and
    the measurement records; results of disjoint seed sets merge."""

    def __init__(self, tlist, e_ops, runs, noise, measurement, seeds, stats, keep_runs,
                 heterodyne, nsc_user):
        self.times = list(tlist)
        self.e_ops = e_ops
        self.seeds = list(seeds)
        self.stats = stats
        self.heterodyne = heterodyne
        vals = []
        for j, O in enumerate(e_ops):
            v = runs[:, :, j]
            if O.isherm:
                v = v.real
            vals.append(v)
        self._sum_expect = [v.sum(axis=0) for v in vals]
        self._sum2_expect = [(v ** 2).sum(axis=0) for v in vals]
        self._set_averages(runs.shape[0])
        self.runs_expect = [v.copy() for v in vals] if keep_runs else None
        # dW summed over every output interval: [ntraj, n_sc, ntimes - 1]
        self._noise = noise
        self._measurement = measurement
        self._nsc_user = nsc_user

    def _set_averages(self, n):
        self.num_trajectories = n
        self.ntraj = n
        self.average_expect = [s / n for s in self._sum_expect]
        self.std_expect = [np.sqrt(np.abs(s2 / n - np.abs((s / n) ** 2)))
                           for s, s2 in zip(self._sum_expect, self._sum2_expect)]
        self.expect = self.average_expect

    @property
    def dW(self):
        """Wiener increments per trajectory and output interval, shape
        (ntraj, n_sc, ntimes - 1), (ntraj, n_sc, 2, ntimes - 1) for
        heterodyne (StochasticTrajResult.dW, stochastic.py:68-81)."""
        if self._noise is None:
            return None
        if self.heterodyne:
            return self._noise.reshape(self._noise.shape[0], -1, 2, self._noise.shape[-1])
        return self._noise

    @property
    def wiener_process(self):
        if self._noise is None:
            return None
        W = np.zeros(self._noise.shape[:2] + (len(self.times),))
        np.cumsum(self._noise, axis=2, out=W[:, :, 1:])
        if self.heterodyne:
            W = W.reshape(W.shape[0], -1, 2, W.shape[-1])
        return W

    @property
    def measurement(self):
        return self._measurement

    def merge(self, other, p=None):
        """p rho + (1 - p) rho' of two results (MultiTrajResult.merge,
        multitrajresult.py:935-1037)."""
        if not isinstance(other, StochasticResult):
            return NotImplemented
        if self.heterodyne != other.heterodyne:
            raise ValueError("Can't merge heterodyne and homodyne results")
        if len(self.e_ops) != len(other.e_ops) or list(self.times) != list(other.times):
            raise ValueError("Shared `e_ops` and `times` are required to merge results")
        n = self.num_trajectories + other.num_trajectories
        p_equal = self.num_trajectories / n
        p = p_equal if p is None else p
        w1, w2 = p / p_equal, (1 - p) / (1 - p_equal)
        new = StochasticResult.__new__(StochasticResult)
        new.times, new.e_ops, new.heterodyne = list(self.times), self.e_ops, self.heterodyne
        new.seeds = self.seeds + other.seeds
        new.stats = dict(self.stats)
        new.stats["end_condition"] = "Merged results"
        new._sum_expect = [w1 * a + w2 * b for a, b in zip(self._sum_expect, other._sum_expect)]
        new._sum2_expect = [w1 * a + w2 * b
                            for a, b in zip(self._sum2_expect, other._sum2_expect)]
        new._set_averages(n)
        cat = (lambda a, b: np.concatenate([a, b]) if a is not None and b is not None else None)
        new.runs_expect = ([np.concatenate([a, b]) for a, b in
                            zip(self.runs_expect, other.runs_expect)]
                           if self.runs_expect is not None and other.runs_expect is not None
                           else None)
        new._noise = cat(self._noise, other._noise)
        new._measurement = cat(self._measurement, other._measurement)
        new._nsc_user = self._nsc_user
        return new

    def __add__(self, other):
        return self.merge(other)

    def __repr__(self):
        return f"<StochasticResult ntraj={self.num_trajectories} ntimes={len(self.times)}>"


class _StochasticSolver:
    """Common part of SSESolver / SMESolver: the system on the device and
    the batched run."""
    _open = False
    name = "stochastic"
    solver_options = {
        "store_final_state": False,
        "store_states": None,
        "keep_runs_results": False,
        "normalize_output": False,
        "method": "platen",
        "store_measurement": "",
        "dt": 0.001,
        "grid": 0,
    }

    def __init__(self, H, sc_ops, heterodyne=False, c_ops=(), options=None):
        H = _qobj(H)
        if not isinstance(H, Qobj) or not H.isoper:
            raise TypeError("The Hamiltonian must be an operator")
        if isinstance(sc_ops, (Qobj, QobjEvo)):
            sc_ops = [sc_ops]
        if isinstance(c_ops, (Qobj, QobjEvo)):
            c_ops = [c_ops]
        sc_user = [_qobj(c) for c in sc_ops]
        c_ops = [_qobj(c) for c in c_ops]
        if not self._open and c_ops:
            raise ValueError("c_ops are not supported by ssesolve.")
        self.heterodyne = bool(heterodyne)
        sc = []
        for c in sc_user:
            sc += [c * (1 / np.sqrt(2)), c * (-1j / np.sqrt(2))] if heterodyne else [c]
        if not 1 <= len(sc) <= MAX_SC:
            raise ValueError(f"1 to {MAX_SC} stochastic operators (after heterodyne splitting)")
        self.H, self.sc_ops, self.c_ops = H, sc, c_ops
        self.d = H.shape[0]
        self.n = self.d * self.d if self._open else self.d
        if self.n > MAX_N:
            raise ValueError(f"the state has {self.n} entries; the trajectory kernel takes "
                             f"at most {MAX_N}")
        if heterodyne:
            self.m_ops = [m for c in sc_user for m in (c + c.dag(), -1j * (c - c.dag()))]
            self.dW_factors = np.ones(2 * len(sc_user)) * 2 ** 0.5
        else:
            self.m_ops = [c + c.dag() for c in sc_user]
            self.dW_factors = np.ones(len(sc_user))
        self._nsc_user = len(sc_user)
        self.options = {**self.solver_options, **(options or {})}
        extra = set(self.options) - set(self.solver_options)
        if extra:
            raise KeyError(f"Options {sorted(extra)} are not supported.")
        if self.options["method"] not in METHODS:
            raise ValueError(f"method must be one of {sorted(METHODS)}")
        self._dev = None  # the device operators, uploaded by the first run

    def _system(self):
        """(L, [M_i]) as CSR: the drift operator and the operators of the
        diffusions (ssystem.pyx:119-250)."""
        d = self.d
        if self._open:
            L = liouvillian(self.H, self.sc_ops + self.c_ops).to_csr()
            eye = sp.identity(d, dtype=np.complex128, format="csr")
            M = [sp.csr_matrix(sp.kron(eye, c.to_csr()) + sp.kron(c.to_csr().conj(), eye))
                 for c in self.sc_ops]
            return L, M
        L = -1j * self.H.to_csr()
        for c in self.sc_ops:
            cc = c.to_csr()
            L = L - 0.5 * (cc.conj().T @ cc)
        return sp.csr_matrix(L), [sp.csr_matrix(c.to_csr()) for c in self.sc_ops]

    def _upload(self):
        L, M = self._system()
        n = self.n
        lw, lcol, lval = _ell(L, n)
        mw = max(_ell(Mi, n)[0] for Mi in M)
        blocks = [_ell(Mi, n, mw) for Mi in M]
        mcol = np.concatenate([b[1] for b in blocks])
        mval = np.concatenate([b[2] for b in blocks])
        self._lw, self._mw = lw, mw
        self._dev = {
            "lcol": _data.DeviceBuffer.from_array(np.ascontiguousarray(lcol)),
            "lval": _data.DeviceBuffer.from_array(np.ascontiguousarray(lval)),
            "mcol": _data.DeviceBuffer.from_array(np.ascontiguousarray(mcol)),
            "mval": _data.DeviceBuffer.from_array(np.ascontiguousarray(mval)),
        }

    def _prepare_state(self, state):
        if self._open:
            rho = ket2dm(state) if state.isket else state
            return rho.full().ravel(order="F")
        if not state.isket:
            raise TypeError("ssesolve needs a ket")
        return state.full().ravel()

    def noise(self, seeds, nsteps):
        """The Wiener increments of every trajectory: [ntraj, nsteps, n_sc]
        (sode/_noise.py:15-30)."""
        nsc = len(self.sc_ops)
        dt = self.options["dt"]
        out = np.empty((len(seeds), nsteps, nsc))
        for i, s in enumerate(seeds):
            out[i] = np.random.default_rng(s).normal(0, np.sqrt(dt),
                                                     size=(nsteps, 1, nsc))[:, 0, :]
        return out

    def run(self, state, tlist, ntraj=1, *, args=None, e_ops=(), seeds=None, timeout=None,
            target_tol=None):
        t0 = time.time()
        if target_tol is not None:
            raise NotImplementedError("target_tol is not supported by the batched kernel")
        tlist = np.asarray(tlist, dtype=np.float64)
        if isinstance(e_ops, (Qobj, QobjEvo)):
            e_ops = [e_ops]
        e_ops = [_qobj(O) for O in (e_ops or [])]
        store_m = self.options["store_measurement"]
        all_ops = e_ops + (self.m_ops if store_m else [])
        if len(all_ops) > MAX_E:
            raise ValueError(f"at most {MAX_E} e_ops (with the measurement operators)")
        for O in all_ops:
            if O.shape != (self.d, self.d):
                raise ValueError(f"incompatible dimensions {O.dims} and {self.H.dims}")
        dt = float(self.options["dt"])
        counts = step_counts(tlist, dt)
        nsteps = int(sum(counts))
        ss = _read_seeds(seeds, ntraj)
        dW = self.noise(ss, nsteps)
        x0 = np.ascontiguousarray(self._prepare_state(state), dtype=np.complex128)
        rp, cl, vl = _csr_block([O.to_csr() for O in all_ops], self.d, self.d)
        ne = len(all_ops)
        nout = len(tlist)
        t_prep = time.time() - t0
        t1 = time.time()
        bufs = {
            "counts": _data.DeviceBuffer.from_array(np.asarray(counts or [0], np.int32)),
            "dW": _data.DeviceBuffer.from_array(np.ascontiguousarray(dW) if nsteps
                                                 else np.zeros(1)),
            "x0": _data.DeviceBuffer.from_array(x0),
            "er": _data.DeviceBuffer.from_array(rp),
            "ec": _data.DeviceBuffer.from_array(cl),
            "ew": _data.DeviceBuffer.from_array(vl),
            "out": _data.DeviceBuffer(16 * max(ntraj * nout * max(ne, 1), 1)),
            "info": _data.DeviceBuffer(4 * max(ntraj, 1)),
        }
        vp = lambda b: C.c_void_p(b.ptr)
        if self._dev is None:
            self._upload()
        D = self._dev
        _lib.call("qa_sde_run", self.n, self.d, int(self._open), len(self.sc_ops),
                  METHODS[self.options["method"]], ntraj, nout, vp(bufs["counts"]), dt,
                  nsteps, self._lw, vp(D["lcol"]), vp(D["lval"]), self._mw, vp(D["mcol"]),
                  vp(D["mval"]), vp(bufs["dW"]), vp(bufs["x0"]), ne, vp(bufs["er"]),
                  vp(bufs["ec"]), vp(bufs["ew"]), vp(bufs["out"]), vp(bufs["info"]),
                  int(self.options["grid"]), None)
        runs = np.empty((ntraj, nout, max(ne, 1)), dtype=np.complex128)
        bufs["out"].download(runs)
        info = np.empty(ntraj, dtype=np.int32)
        bufs["info"].download(info)
        t_run = time.time() - t1
        if (info != 0).any():
            bad = np.flatnonzero(info)
            raise RuntimeError(f"{len(bad)} trajectories stopped being finite (first: "
                               f"{bad[0]}); use a smaller dt")
        # the increments summed over every output interval
        edges = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        noise = np.stack([dW[:, a:b, :].sum(axis=1) for a, b in zip(edges[:-1], edges[1:])],
                         axis=2) if nout > 1 else np.zeros((ntraj, len(self.sc_ops), 0))
        measurement = None
        if store_m:
            m_expect = runs[:, :, len(e_ops):len(e_ops) + len(self.m_ops)].real
            m_expect = np.transpose(m_expect, (0, 2, 1))  # [ntraj, n_m, nout]
            if store_m == "start":
                m = m_expect[:, :, :-1]
            elif store_m == "middle":
                m = 0.5 * (m_expect[:, :, 1:] + m_expect[:, :, :-1])
            elif store_m in ("end", True):
                m = m_expect[:, :, 1:]
            else:
                raise ValueError("store_measurement must be in {'start', 'middle', 'end', ''}")
            m = m + self.dW_factors[None, :, None] * noise / np.diff(tlist)[None, None, :]
            measurement = (m.reshape(ntraj, -1, 2, m.shape[-1]) if self.heterodyne else m)
        stats = {"method": self.options["method"], "preparation time": t_prep,
                 "run time": t_run, "num_steps": nsteps, "solver": self.name}
        return StochasticResult(tlist, e_ops, runs[:, :, :len(e_ops)], noise, measurement, ss,
                                stats, bool(self.options["keep_runs_results"]),
                                self.heterodyne, self._nsc_user)


class SSESolver(_StochasticSolver):
    """Stochastic Schroedinger equation (stochastic.py: SSESolver)."""
    _open = False
    name = "ssesolve"

    def __init__(self, H, sc_ops, heterodyne=False, *, options=None):
        super().__init__(H, sc_ops, heterodyne, (), options)


class SMESolver(_StochasticSolver):
    """Stochastic master equation (stochastic.py: SMESolver)."""
    _open = True
    name = "smesolve"

    def __init__(self, H, sc_ops, heterodyne=False, *, c_ops=(), options=None):
        super().__init__(H, sc_ops, heterodyne, c_ops, options)


def ssesolve(H, psi0, tlist, sc_ops=(), heterodyne=False, *, e_ops=(), args=None, ntraj=500,
             options=None, seeds=None, target_tol=None, timeout=None):
    """stochastic.py:455-585."""
    return SSESolver(H, sc_ops, heterodyne, options=options).run(
        psi0, tlist, ntraj, e_ops=e_ops, seeds=seeds, target_tol=target_tol)


def smesolve(H, rho0, tlist, c_ops=(), sc_ops=(), heterodyne=False, *, e_ops=(), args=None,
             ntraj=500, options=None, seeds=None, target_tol=None, timeout=None):
    """stochastic.py:311-452."""
    return SMESolver(H, sc_ops, heterodyne, c_ops=c_ops, options=options).run(
        rho0, tlist, ntraj, e_ops=e_ops, seeds=seeds, target_tol=target_tol)
