"""mcsolve / MCSolver: batched Monte-Carlo wave-function trajectories.

Same physics and options as qutip/solver/mcsolve.py (MCSolver :424-821,
MCIntegrator :228-419): H_eff = -iH - 1/2 sum_k c_k^dag c_k (:491-496),
norm-decay waiting times, log-ratio secant search of the collapse time,
channel choice weighted by <psi|c_k^dag c_k|psi>.  The trajectories, instead
of being farmed to CPU processes one by one (qutip/solver/parallel.py), all
run inside one persistent kernel (``qa_mc_run``): each workgroup integrates
whole trajectories with its own step sizes and jump times.

Random numbers follow the reference stream exactly: trajectory i uses
``numpy.random.default_rng(seeds[i])`` (PCG64; qutip/solver/multitraj.py:
354-393) and consumes one draw for the first target norm, then per jump one
draw for the channel (when there is more than one) and one for the next
target (mcsolve.py:276-279, 388, 405).  The draws are made here, on the
host, before the launch and streamed to the device.

The ensemble reduction is MultiTrajResult's: average = sum / ntraj,
std = sqrt(|mean(x^2) - |mean(x)|^2|) (qutip/solver/multitrajresult.py:
261-279, 1116-1124).
"""
import ctypes as C
import time

import numpy as np
import scipy.sparse as sp
from numpy.random import SeedSequence

from .. import _lib
from .. import data as _data
from ..core import Qobj, QobjEvo
from .integrator import IntegratorException
from .tableaus import TABLEAUS

__all__ = ["mcsolve", "MCSolver", "McResult", "MC_STATUS"]

MC_STATUS = {
    0: "ok",
    -1: ("Too much work done in one call. Try to increase the nsteps parameter "
         "or increasing the tolerance."),
    -2: "Step size becomes too small. Try increasing tolerance.",
    -3: "Step outside available range.",
    -10: ("Could not find the collapse time within desired tolerance. Increase "
          "accuracy of the ODE solver or lower the tolerance with the options "
          "'norm_steps', 'norm_tol', 'norm_t_tol'."),
    -11: "pre-drawn random numbers exhausted",
    -12: "no progress",
}


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _vp(p):
    return C.c_void_p(p)


class McResult:
    """Ensemble result with the attributes of qutip's McResult used here.
    The per-time sums of <O> and <O>^2 over the trajectories are kept
    (_TrajectorySum, multitrajresult.py:1062-1175), so results of disjoint
    seed sets can be merged."""

    def __init__(self, tlist, e_ops, runs_expect, info, col_times, col_which, seeds, stats,
                 keep_runs):
        self.times = list(tlist)
        self.e_ops = e_ops
        self.seeds = list(seeds) if seeds is not None else []
        self.stats = stats
        n = runs_expect.shape[0]
        vals = []
        for j, O in enumerate(e_ops):
            v = runs_expect[:, :, j]
            if O.isherm:
                v = v.real
            vals.append(v)
        self._sum_expect = [v.sum(axis=0) for v in vals]
        self._sum2_expect = [(v ** 2).sum(axis=0) for v in vals]
        self._set_averages(n)
        self.runs_expect = [v.copy() for v in vals] if keep_runs else None
        nj = info[:, 1]
        self.col_times = [list(col_times[i, :min(nj[i], col_times.shape[1])]) for i in range(n)]
        self.col_which = [list(col_which[i, :min(nj[i], col_which.shape[1])]) for i in range(n)]
        self.num_jumps = nj.copy()

    def _set_averages(self, n):
        """Mean and standard deviation from the sums (multitrajresult.py:
        261-279): std = sqrt(|mean(x^2) - |mean(x)|^2|)."""
        self.num_trajectories = n
        self.ntraj = n
        self.average_expect = [s / n for s in self._sum_expect]
        self.std_expect = [np.sqrt(np.abs(s2 / n - np.abs((s / n) ** 2)))
                           for s, s2 in zip(self._sum_expect, self._sum2_expect)]
        self.expect = self.average_expect

    @property
    def average_e_data(self):
        return dict(enumerate(self.average_expect))

    def merge(self, other, p=None):
        """The ensemble p rho + (1 - p) rho' of two results on the same
        e_ops and times; p defaults to N / (N + N'), every trajectory
        weighing the same (MultiTrajResult.merge, multitrajresult.py:
        935-1037: the sums are combined with weights p / p_equal and
        (1 - p) / (1 - p_equal))."""
        if not isinstance(other, McResult):
            return NotImplemented
        if len(self.e_ops) != len(other.e_ops) or not all(
                a is b or np.array_equal(a.full(), b.full())
                for a, b in zip(self.e_ops, other.e_ops)):
            raise ValueError("Shared `e_ops` is required to merge results")
        if list(self.times) != list(other.times):
            raise ValueError("Shared `times` are is required to merge results")
        n = self.num_trajectories + other.num_trajectories
        p_equal = self.num_trajectories / n
        if p is None:
            p = p_equal
        w1, w2 = p / p_equal, (1 - p) / (1 - p_equal)
        new = McResult.__new__(McResult)
        new.times = list(self.times)
        new.e_ops = self.e_ops
        new.seeds = self.seeds + other.seeds
        new.stats = dict(self.stats)
        new.stats["run time"] = self.stats.get("run time", 0.0) + other.stats.get("run time", 0.0)
        new.stats["end_condition"] = "Merged results"
        new._sum_expect = [w1 * a + w2 * b for a, b in zip(self._sum_expect, other._sum_expect)]
        new._sum2_expect = [w1 * a + w2 * b
                            for a, b in zip(self._sum2_expect, other._sum2_expect)]
        new._set_averages(n)
        new.runs_expect = ([np.concatenate([a, b]) for a, b in
                            zip(self.runs_expect, other.runs_expect)]
                           if self.runs_expect is not None and other.runs_expect is not None
                           else None)
        new.col_times = self.col_times + other.col_times
        new.col_which = self.col_which + other.col_which
        new.num_jumps = np.concatenate([self.num_jumps, other.num_jumps])
        return new

    def __add__(self, other):
        return self.merge(other, p=None)

    def __repr__(self):
        return f"<McResult ntraj={self.num_trajectories} ntimes={len(self.times)}>"


def _read_seeds(seeds, ntraj):
    """MultiTrajSolver._read_seed (qutip/solver/multitraj.py:354-374)."""
    if seeds is None:
        return SeedSequence().spawn(ntraj)
    if isinstance(seeds, SeedSequence):
        return seeds.spawn(ntraj)
    if not isinstance(seeds, list):
        return SeedSequence(seeds).spawn(ntraj)
    if len(seeds) >= ntraj:
        return [s if isinstance(s, SeedSequence) else SeedSequence(s) for s in seeds[:ntraj]]
    raise ValueError("A seed list must be longer than ntraj")


def draws_for(seeds, n):
    """The first n uniform draws of default_rng(seed) for every seed."""
    out = np.empty((len(seeds), n), dtype=np.float64)
    for i, s in enumerate(seeds):
        out[i] = np.random.default_rng(s).random(n)
    return out


def _csr_block(ops, d, fp32):
    """Concatenate CSR operators: absolute row pointers, one block of d + 1 per
    operator."""
    rowptr, cols, vals = [], [], []
    off = 0
    for O in ops:
        A = sp.csr_matrix(O, dtype=np.complex128)
        A.sort_indices()
        if A.shape != (d, d):
            raise ValueError(f"operator of shape {A.shape}, expected {(d, d)}")
        rowptr.append(A.indptr.astype(np.int64) + off)
        cols.append(A.indices)
        vals.append(A.data)
        off += A.nnz
    rp = np.concatenate(rowptr).astype(np.int32) if rowptr else np.zeros(1, np.int32)
    cl = np.concatenate(cols).astype(np.int32) if off else np.zeros(1, np.int32)
    vl = np.concatenate(vals) if off else np.zeros(1, np.complex128)
    vl = vl.astype(np.complex64 if fp32 else np.complex128)
    return (_data.DeviceBuffer.from_array(rp), _data.DeviceBuffer.from_array(cl),
            _data.DeviceBuffer.from_array(vl))


def _ell_host(A, d, fp32):
    """Column-major ELL of a d x d operator for the trajectory kernel: width W
    = the longest row rounded up to a multiple of 4; entry u of row r at
    [u, r]; shorter rows padded with value 0 and their own row index as the
    column.  Returns (W, col [max(W,1), d] int32, val [max(W,1), d])."""
    A = sp.csr_matrix(A, dtype=np.complex128)
    A.sort_indices()
    if A.shape != (d, d):
        raise ValueError(f"operator of shape {A.shape}, expected {(d, d)}")
    lens = np.diff(A.indptr)
    W = -(-int(lens.max()) // 4) * 4 if d else 0
    col = np.tile(np.arange(d, dtype=np.int32), (max(W, 1), 1))
    val = np.zeros((max(W, 1), d), dtype=np.complex128)
    rows = np.repeat(np.arange(d), lens)
    slot = np.arange(A.nnz) - np.repeat(A.indptr[:-1], lens)
    col[slot, rows] = A.indices
    val[slot, rows] = A.data
    return W, col, val.astype(np.complex64 if fp32 else np.complex128)


def _ell_block(A, d, fp32):
    """_ell_host uploaded: (W, column buffer, value buffer)."""
    W, col, val = _ell_host(A, d, fp32)
    return (W, _data.DeviceBuffer.from_array(np.ascontiguousarray(col)),
            _data.DeviceBuffer.from_array(np.ascontiguousarray(val)))


class MCSolver:
    """Monte-Carlo wave-function solver, all trajectories in one persistent
    kernel.  ``options`` keys as qutip's MCSolver.solver_options
    (mcsolve.py:449-466) and the integrator options of IntegratorVern7, plus
    ``precision`` ('fp64' | 'fp32': complex64 state and RHS with FP64 norms),
    ``max_draws`` / ``max_jumps`` (buffer sizes, grown automatically)."""
    name = "mcsolve"
    solver_options = {
        "progress_bar": "",
        "progress_kwargs": {},
        "store_final_state": False,
        "store_states": None,
        "keep_runs_results": False,
        "map": "serial",
        "mpi_options": {},
        "num_cpus": None,
        "bitgenerator": None,
        "method": "vern7",
        "mc_corr_eps": 1e-10,
        "norm_steps": 25,
        "norm_t_tol": 1e-6,
        "norm_tol": 1e-4,
        "norm_min_step": 0.1,
        "improved_sampling": False,
        # integrator (IntegratorVern7.integrator_options)
        "atol": 1e-8,
        "rtol": 1e-6,
        "nsteps": 1000,
        "first_step": 0,
        "max_step": 0,
        "min_step": 0,
        "interpolate": True,
        # device
        "precision": "fp64",
        "max_draws": 64,
        "max_jumps": 64,
        "grid": 0,
    }

    def __init__(self, H, c_ops, *, options=None):
        if isinstance(H, QobjEvo):
            H = H.qobj
        if isinstance(c_ops, (Qobj, QobjEvo)):
            c_ops = [c_ops]
        c_ops = [c.qobj if isinstance(c, QobjEvo) else c for c in c_ops]
        if not c_ops:
            raise ValueError("mcsolve needs at least one collapse operator")
        if H.issuper:
            raise NotImplementedError("a Liouvillian H is not supported by the batched kernel")
        self.H = H
        self.c_ops = c_ops
        self.options = dict(self.solver_options)
        for k, v in (options or {}).items():
            if k not in self.solver_options:
                raise KeyError(f"Options ['{k}'] are not supported.")
            self.options[k] = v
        if self.options["improved_sampling"]:
            raise NotImplementedError("improved_sampling is not supported by the batched kernel")
        if self.options["method"] not in TABLEAUS:
            raise ValueError(f"method must be one of {sorted(TABLEAUS)}")
        self.d = H.shape[0]
        Hc = H.to_csr()
        Heff = -1j * Hc
        self._n_ops = []
        for c in c_ops:
            cc = c.to_csr()
            if cc.nnz > 0.05 * self.d * self.d:
                # a dense collapse operator: the product by BLAS (a sparse
                # product of dense factors costs orders of magnitude more)
                cd = cc.toarray()
                n = sp.csr_matrix(cd.conj().T @ cd)
            else:
                n = (cc.conj().T @ cc).tocsr()
            self._n_ops.append(n)
            Heff = Heff - 0.5 * n
        Heff = sp.csr_matrix(Heff)
        Heff.sum_duplicates()
        Heff.eliminate_zeros()
        Heff.sort_indices()
        self.H_eff = Heff
        self._dev = None

    def _device_ops(self, fp32):
        if self._dev is None or self._dev[0] != fp32:
            h = _ell_block(self.H_eff, self.d, fp32)
            c = _csr_block([c.to_csr() for c in self.c_ops], self.d, fp32)
            tab = TABLEAUS[self.options["method"]]
            size = _lib.lib().qa_mc_tableau_size()
            buf = np.zeros(size, dtype=np.uint8)
            a, b, cc = (np.ascontiguousarray(tab[k], dtype=np.float64) for k in ("a", "b", "c"))
            e = np.ascontiguousarray(tab["e"], dtype=np.float64)
            bi = np.ascontiguousarray(tab["bi"], dtype=np.float64)
            _lib.call("qa_mc_pack_tableau", buf.ctypes.data_as(C.c_void_p), int(tab["order"]),
                      len(b), len(cc), bi.shape[1], _dp(a), _dp(b), _dp(cc), _dp(e), _dp(bi),
                      int(bool(self.options["interpolate"])))
            tab_dev = _data.DeviceBuffer.from_array(buf)
            self._dev = (fp32, h, c, tab_dev)
        return self._dev

    def run_batch(self, psi0, tlist, e_ops, draws, max_jumps=64, fp32=None):
        """Device part of run(): trajectories from pre-drawn uniform numbers
        ``draws`` (ntraj x max_draws, numpy or DeviceBuffer) -> (runs_expect,
        info, col_times, col_which) as numpy arrays."""
        if fp32 is None:
            fp32 = self.options["precision"] == "fp32"
        return self._launch(np.asarray(psi0).ravel(), tlist, e_ops, draws, max_jumps, fp32)

    def launch_async(self, psi0, tlist, e_ops, draws, max_jumps=64, stream=None, fp32=None):
        """Enqueue the trajectory kernel on ``stream`` and return a handle that
        keeps the device buffers alive; ``collect_async`` reads the results
        once the stream has reached this point."""
        o = self.options
        if fp32 is None:
            fp32 = o["precision"] == "fp32"
        if isinstance(draws, tuple):  # (DeviceBuffer, ntraj, max_draws)
            dr, ntraj, max_draws = draws
        else:
            ntraj, max_draws = draws.shape
            dr = _data.DeviceBuffer.from_array(np.ascontiguousarray(draws))
        ntimes = len(tlist)
        _, h, c, tab = self._device_ops(fp32)
        e = _csr_block([O.to_csr() for O in e_ops] or [sp.identity(self.d)], self.d, fp32)
        neops = len(e_ops)
        H = {"e": e, "dr": dr, "ntraj": ntraj, "ntimes": ntimes, "neops": neops,
             "max_jumps": max_jumps, "stream": stream}
        H["times"] = _data.DeviceBuffer.from_array(np.asarray(tlist, dtype=np.float64))
        H["p0"] = _data.DeviceBuffer.from_array(np.ascontiguousarray(np.asarray(psi0).ravel(),
                                                                     dtype=np.complex128))
        H["ex"] = _data.DeviceBuffer(16 * ntraj * ntimes * max(neops, 1))
        H["ct"] = _data.DeviceBuffer(8 * ntraj * max_jumps)
        H["cw"] = _data.DeviceBuffer(4 * ntraj * max_jumps)
        H["info"] = _data.DeviceBuffer(16 * ntraj)
        L = _lib.lib()
        rc = L.qa_mc_run(
            self.d, ntraj, ntimes, _vp(H["times"].ptr), _vp(H["p0"].ptr), int(h[0]),
            _vp(h[1].ptr), _vp(h[2].ptr), len(self.c_ops), _vp(c[0].ptr), _vp(c[1].ptr),
            _vp(c[2].ptr), neops, _vp(e[0].ptr), _vp(e[1].ptr), _vp(e[2].ptr), _vp(tab.ptr),
            _vp(dr.ptr), max_draws, float(o["rtol"]), float(o["atol"]), int(o["nsteps"]),
            float(o["first_step"]), float(o["min_step"]), float(o["max_step"]),
            int(o["norm_steps"]), float(o["norm_t_tol"]), float(o["norm_tol"]),
            float(o["norm_min_step"]), float(o["mc_corr_eps"]), int(fp32), _vp(H["ex"].ptr),
            _vp(H["ct"].ptr), _vp(H["cw"].ptr), max_jumps, _vp(H["info"].ptr), int(o["grid"]),
            stream)
        _lib.check(rc, "qa_mc_run")
        return H

    def collect_async(self, H, with_jumps=False):
        """Results of a finished launch_async: (runs_expect, info[, col_times,
        col_which])."""
        ntraj, ntimes, neops, mj = H["ntraj"], H["ntimes"], H["neops"], H["max_jumps"]
        if H["stream"] is not None:
            _lib.call("qa_stream_sync", H["stream"])
        out = np.empty((ntraj, ntimes, max(neops, 1)), dtype=np.complex128)
        H["ex"].download(out)
        inf = np.empty((ntraj, 4), dtype=np.int32)
        H["info"].download(inf)
        if not with_jumps:
            return out[:, :, :neops], inf
        cts = np.empty((ntraj, mj), dtype=np.float64)
        H["ct"].download(cts)
        cws = np.empty((ntraj, mj), dtype=np.int32)
        H["cw"].download(cws)
        return out[:, :, :neops], inf, cts, cws

    def _launch(self, psi0, tlist, e_ops, draws, max_jumps, fp32):
        H = self.launch_async(psi0, tlist, e_ops, draws, max_jumps, None, fp32)
        _lib.call("qa_device_sync")
        return self.collect_async(H, with_jumps=True)

    def run_plain(self, psi0, tlist, e_ops, draws):
        """The plain drop-in form of the trajectory loop: what qutip's
        MCSolver does with method='hip_rk' (MCIntegrator, mcsolve.py:245-406,
        over the Integrator protocol).  Trajectories run one after another;
        every step is an IntegratorHipRK.mcstep on the fused engine of
        H_eff, and the norms, the collapse-time search (log-ratio secant,
        :321-369) and the jumps (:371-406) run on device data through the
        data layer, with host round trips in between.  ``draws``: ntraj x
        max_draws uniform numbers, consumed in the reference's order (the
        numbers the batched kernel reads).  Returns (runs_expect [ntraj,
        ntimes, n_e_ops], jumps [ntraj])."""
        from .hip_rk import IntegratorHipRK
        o = self.options
        integ = IntegratorHipRK(QobjEvo(Qobj(self.H_eff)).matmul_data, {
            k: o[k] for k in ("rtol", "atol", "nsteps", "first_step", "min_step", "max_step",
                              "interpolate")} | {"tableau": o["method"]})
        cs = [_data.HipCSR(c.to_csr()) for c in self.c_ops]
        ns = [_data.HipCSR(n) for n in self._n_ops]
        eo = [_data.HipCSR(O.to_csr()) for O in e_ops]
        tlist = np.asarray(tlist, dtype=np.float64)
        draws = np.asarray(draws, dtype=np.float64)
        ntraj = draws.shape[0]
        runs = np.zeros((ntraj, len(tlist), len(eo)), dtype=np.complex128)
        jumps = np.zeros(ntraj, dtype=np.int64)
        prob = lambda y: _data.frobenius(y) ** 2
        p0 = _data.HipDense.from_array(np.asarray(psi0, dtype=np.complex128).reshape(-1, 1))
        for k in range(ntraj):
            it = iter(draws[k])
            target = next(it)
            integ.set_state(tlist[0], p0)
            nrm0 = np.sqrt(prob(p0))
            runs[k, 0] = [_data.expect(O, p0) / nrm0 ** 2 for O in eo]
            for i, t in enumerate(tlist[1:], 1):
                t_old, y_old = integ.get_state(copy=True)
                norm_old = prob(y_old)
                while t_old < t:
                    t_step, state = integ.mcstep(t, copy=True)
                    nrm = prob(state)
                    if nrm <= target:
                        t_col, state = self._plain_collapse_time(integ, prob, target, norm_old,
                                                                 nrm, t_old, t_step)
                        target = self._plain_collapse(integ, cs, ns, t_col, state, it, target,
                                                      jumps, k)
                        t_old, y_old = integ.get_state(copy=True)
                        norm_old = 1.0
                    else:
                        t_old, y_old, norm_old = t_step, state, nrm
                n2 = prob(y_old)
                runs[k, i] = [_data.expect(O, y_old) / n2 for O in eo]
        return runs, jumps

    def _plain_collapse_time(self, integ, prob, target, norm_old, norm, t_prev, t_final):
        """_find_collapse_time (mcsolve.py:321-369) through mcstep."""
        o = self.options
        tries = 0
        cutoff = o["norm_min_step"]
        while tries < o["norm_steps"]:
            tries += 1
            if (t_final - t_prev) < o["norm_t_tol"]:
                t_guess = t_final
                _, state = integ.get_state(copy=True)
                break
            ratio = np.log(norm_old / target) / np.log(norm_old / norm)
            ratio = min(max(ratio, cutoff), 1 - cutoff)
            t_guess = t_prev + (t_final - t_prev) * ratio
            if (t_guess - t_prev) < o["norm_t_tol"]:
                t_guess = t_prev + o["norm_t_tol"]
            _, state = integ.mcstep(t_guess, copy=True)
            n_guess = prob(state)
            if abs(target - n_guess) < o["norm_tol"] * target:
                break
            if n_guess < target:
                t_final, norm = t_guess, n_guess
            else:
                t_prev, norm_old = t_guess, n_guess
        if tries >= o["norm_steps"]:
            raise RuntimeError(MC_STATUS[-10])
        return t_guess, state

    def _plain_collapse(self, integ, cs, ns, t, state, draws, target, jumps, k):
        """_do_collapse (mcsolve.py:371-406): the channel by <c^dag c>, the
        jump, a new target norm; returns the target."""
        which = 0
        if len(ns) > 1:
            probs = [_data.inner(state, _data.matmul(n, state)).real for n in ns]
            left = sum(probs) * next(draws) - probs[0]
            while left > 0 and which < len(ns) - 1:
                which += 1
                left -= probs[which]
        new = _data.matmul(cs[which], state)
        nn = _data.frobenius(new)
        if nn < self.options["mc_corr_eps"]:
            new = _data.mul(state, 1.0 / _data.frobenius(state))
        else:
            new = _data.mul(new, 1.0 / nn)
            jumps[k] += 1
            target = next(draws)
        integ.set_state(t, new)
        return target

    def run(self, state, tlist, ntraj=1, *, args=None, e_ops=None, timeout=None,
            target_tol=None, seeds=None):
        t0 = time.time()
        if target_tol is not None:
            raise NotImplementedError("target_tol is not supported by the batched kernel")
        if isinstance(e_ops, dict):
            e_ops = list(e_ops.values())
        e_ops = [] if e_ops is None else ([e_ops] if isinstance(e_ops, Qobj) else list(e_ops))
        e_ops = [O(0.0) if isinstance(O, QobjEvo) else O for O in e_ops]
        if not state.isket:
            raise NotImplementedError("the batched kernel runs ket initial states")
        tlist = np.asarray(tlist, dtype=np.float64)
        ss = _read_seeds(seeds, ntraj)
        fp32 = self.options["precision"] == "fp32"
        psi0 = state.full().ravel()
        max_draws = int(self.options["max_draws"])
        max_jumps = int(self.options["max_jumps"])
        runs = np.zeros((ntraj, len(tlist), len(e_ops)), dtype=np.complex128)
        info = np.zeros((ntraj, 4), dtype=np.int32)
        cts = np.zeros((ntraj, max_jumps))
        cws = np.zeros((ntraj, max_jumps), dtype=np.int32)
        todo = np.arange(ntraj)
        t_prep = time.time() - t0
        t1 = time.time()
        launches = 0
        while len(todo):
            draws = draws_for([ss[i] for i in todo], max_draws)
            ex, inf, ct, cw = self._launch(psi0, tlist, e_ops, draws, max_jumps, fp32)
            launches += 1
            if ct.shape[1] > cts.shape[1]:
                cts = np.pad(cts, ((0, 0), (0, ct.shape[1] - cts.shape[1])))
                cws = np.pad(cws, ((0, 0), (0, cw.shape[1] - cws.shape[1])))
            runs[todo] = ex
            info[todo] = inf
            cts[todo, :ct.shape[1]] = ct
            cws[todo, :cw.shape[1]] = cw
            bad = inf[:, 0]
            for code in np.unique(bad[(bad < 0) & (bad != -11)]):
                msg = MC_STATUS.get(int(code), f"status {code}")
                if code == -10:
                    raise RuntimeError(msg)
                raise IntegratorException(msg)
            short = (bad == -11) | (inf[:, 1] > max_jumps)
            todo = todo[short]
            if len(todo):
                max_draws *= 4
                max_jumps = max(max_jumps, int(inf[short, 1].max()) + 8)
                if max_draws > 1 << 20:
                    raise RuntimeError("trajectories need more than 2**20 random draws")
        stats = {"method": f"mcsolve hip ({self.options['method']}, "
                           f"{'complex64' if fp32 else 'complex128'})",
                 "preparation time": t_prep, "run time": time.time() - t1,
                 "launches": launches, "rk_attempts": int(info[:, 3].sum()),
                 "jumps": int(info[:, 1].sum()), "num_collapse": len(self.c_ops),
                 "end_condition": "ntraj reached"}
        return McResult(tlist, e_ops, runs, info, cts, cws, ss, stats,
                        bool(self.options["keep_runs_results"]))


def mcsolve(H, state, tlist, c_ops=(), e_ops=None, ntraj=500, *, args=None, options=None,
            seeds=None, target_tol=None, timeout=None, **kwargs):
    """Monte-Carlo evolution (qutip/solver/mcsolve.py:24-197)."""
    solver = MCSolver(H, list(c_ops), options=options)
    return solver.run(state, tlist, ntraj, e_ops=e_ops, seeds=seeds, target_tol=target_tol)
