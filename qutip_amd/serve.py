"""Continuous-batching solve server: many sesolve / mesolve / mcsolve
requests of different sizes in flight on one GPU.

Requests are admitted while device memory (an HBM budget) and slots allow;
every admitted request owns its device-resident operators, state and work
vectors and runs on its own HIP stream, so the work of concurrent requests
overlaps on the device:

* sesolve: the whole solve in one workgroup (d <= 1024: steps, step
  control, dense output and the expectation values at every output time in
  one kernel), else the fused RK engine on the sparse -iH;
* mesolve: the fused RK engine on the matrix-form Lindblad RHS (MFMA
  products; dense collapse operators make the superoperator infeasible:
  d^4 non-zeros);
* mcsolve: one persistent trajectory kernel for all trajectories of the
  request (complex128; complex64 with FP64 norms on request).

The host loop never waits on one request: it polls each request's progress
(qa_rk_advance / an event after the trajectory kernel), tops up its queued
device work, retires finished requests and admits new arrivals as memory
and slots free up.  Latency is completion minus arrival.

A request's result does not depend on what else is in flight: each request
computes on its own buffers with kernels whose tiling and reduction order
depend only on its own sizes, and there are no cross-request atomics.
"""
import ctypes as C
import time

import numpy as np
import scipy.sparse as sp

from . import _lib
from . import data as _data
from .core import Qobj
from .solver import small as _small
from .solver.hip_rk import RKEngine
from .solver.mcsolve import MCSolver, _read_seeds, draws_for

__all__ = ["SolveRequest", "SolveServer"]


class SolveRequest:
    def __init__(self, kind, H, psi0, tlist, c_ops=(), e_ops=(), ntraj=0, seed=0, arrival=0.0,
                 rid=0, options=None):
        if kind not in ("se", "me", "mc"):
            raise ValueError("kind must be 'se', 'me' or 'mc'")
        self.kind = kind
        self.H = H
        self.psi0 = psi0
        self.tlist = np.asarray(tlist, dtype=np.float64)
        self.c_ops = list(c_ops)
        self.e_ops = list(e_ops)
        self.ntraj = int(ntraj)
        self.seed = seed
        self.arrival = float(arrival)
        self.id = rid
        self.options = {"rtol": 1e-6, "atol": 1e-8, "nsteps": 1000, **(options or {})}
        # filled by the server
        self.result = None
        self.stderr = None
        self.t_admit = None
        self.t_done = None
        self.error = None

    @property
    def d(self):
        return self.H.shape[0]

    def device_bytes(self):
        """Estimate of the request's HBM footprint (admission control)."""
        d = self.d
        if self.kind == "se":
            nnz = self.H.to_csr().nnz
            return 22 * d * 16 + nnz * 24 + (1 << 20)
        if self.kind == "me":
            n = d * d
            return 22 * n * 16 + (2 + 3 * len(self.c_ops)) * n * 8 + (1 << 20)
        return (d * d * (1 + len(self.c_ops)) * 20 + self.ntraj * len(self.tlist) * 32
                + self.ntraj * 1024 + (1 << 20))

    def latency(self):
        return None if self.t_done is None else self.t_done - self.arrival

    def work_estimate(self):
        """Relative device work of the request, for shortest-first
        admission: flops of one RHS times the number of states it carries."""
        d = self.d
        if self.kind == "se":
            return 8.0 * self.H.to_csr().nnz
        if self.kind == "me":
            return (2 + 2 * len(self.c_ops)) * 8.0 * d ** 3
        return 8.0 * d * d * max(self.ntraj, 1)


class _Running:
    """Device-side state of one admitted request."""

    def __init__(self, req):
        self.req = req
        self.done = False
        self.kind = req.kind
        o = req.options
        tl = req.tlist
        if req.kind == "se" and req.d <= _small.MAX_N:
            # the whole solve in one workgroup, on the request's own stream
            A = -1j * req.H.to_csr()
            self.small = _small.SmallSolve(A, "vern7", o["rtol"], o["atol"], o["nsteps"])
            psi = req.psi0.full().ravel()
            kinds, fptr, rows, cols, w = self._quadratic(req, psi.size)
            cat = lambda xs, dt: np.concatenate(xs).astype(dt)
            self.small.set_functionals(np.asarray(kinds, np.int32), np.asarray(fptr, np.int32),
                                       cat(rows, np.int32), cat(cols, np.int32),
                                       cat(w, np.complex128))
            self._first = self._functionals_host(psi, req)
            self.kind = "se_small"
            self.stream = C.c_void_p()
            _lib.call("qa_stream_create", C.byref(self.stream))
            self.event = C.c_void_p()
            _lib.call("qa_event_create", C.byref(self.event))
            self._h = self.small.launch(tl[0], psi, tl[1:], stream=self.stream)
            _lib.call("qa_event_record", self.event, self.stream)
        elif req.kind in ("se", "me"):
            if req.kind == "se":
                A = -1j * req.H.to_csr()
                self.eng = RKEngine(A, "vern7", o["rtol"], o["atol"], o["nsteps"])
                psi = req.psi0.full().ravel()
                kinds, fptr, rows, cols, w = self._quadratic(req, psi.size)
                y0 = psi
            else:
                cs = [c.full() for c in req.c_ops]
                Hnh = req.H.full() - 0.5j * sum(c.conj().T @ c for c in cs)
                self.eng = RKEngine(None, "vern7", o["rtol"], o["atol"], o["nsteps"],
                                    dense=(Hnh, cs), dense_precision=o.get("dense_precision", "fp32"))
                # rho0 = |psi0><psi0| is Hermitian: the upper tiles and their
                # mirror (mesolve's assume_hermitian_state)
                self.eng.set_hermitian(True)
                psi = req.psi0.full().ravel()
                rho = np.outer(psi, psi.conj())
                d = rho.shape[0]
                kinds, fptr, rows, cols, w = [], [0], [], [], []
                for O in req.e_ops + [None]:
                    M = sp.coo_matrix(O.to_csr()) if O is not None else sp.identity(d, format="coo")
                    kinds.append(0)
                    rows.append(np.zeros(M.nnz, dtype=np.int64))
                    cols.append(M.row * d + M.col)
                    w.append(M.data.astype(np.complex128))
                    fptr.append(fptr[-1] + M.nnz)
                y0 = rho.ravel(order="F")
            cat = lambda xs, dt: np.concatenate(xs).astype(dt)
            self.eng.set_functionals(np.asarray(kinds, np.int32), np.asarray(fptr, np.int32),
                                     cat(rows, np.int32), cat(cols, np.int32),
                                     cat(w, np.complex128))
            self._y0 = _data.HipDense.from_array(y0)
            self.eng.set_state(tl[0], self._y0)
            # the value at t0 comes from the initial state (host, once)
            self._first = self._functionals_host(y0, req)
            times = np.ascontiguousarray(tl[1:])
            self._times = times
            _lib.call("qa_rk_begin", self.eng.handle, len(times),
                      times.ctypes.data_as(C.POINTER(C.c_double)), 0, None)
        else:
            # complex128 unless the request asks for complex64: at the
            # mix's tolerances (rtol 1e-8) the complex64 trajectory kernel
            # ran 2-3x slower per request (its error estimate carries FP32
            # rounding of about 1e-7, which a 1e-8 tolerance turns into
            # smaller steps)
            self.solver = MCSolver(req.H, req.c_ops, options={
                "rtol": o["rtol"], "atol": o["atol"], "nsteps": o["nsteps"],
                "precision": o.get("precision", "fp64")})
            seeds = _read_seeds(req.seed, req.ntraj)
            self._draws = draws_for(seeds, 64)
            self.stream = C.c_void_p()
            _lib.call("qa_stream_create", C.byref(self.stream))
            self.event = C.c_void_p()
            _lib.call("qa_event_create", C.byref(self.event))
            self._mc = self.solver.launch_async(req.psi0.full().ravel(), tl, req.e_ops,
                                                self._draws, 64, self.stream)
            _lib.call("qa_event_record", self.event, self.stream)

    @staticmethod
    def _quadratic(req, n):
        """<psi|O|psi> for every e_op and ||psi||^2 as quadratic functionals."""
        kinds, fptr, rows, cols, w = [], [0], [], [], []
        for O in req.e_ops + [None]:
            M = sp.coo_matrix(O.to_csr()) if O is not None else sp.identity(n, format="coo")
            kinds.append(1)
            rows.append(M.row)
            cols.append(M.col)
            w.append(M.data.astype(np.complex128))
            fptr.append(fptr[-1] + M.nnz)
        return kinds, fptr, rows, cols, w

    @staticmethod
    def _functionals_host(y0, req):
        vals = []
        if req.kind == "se":
            nrm = np.vdot(y0, y0).real
            for O in req.e_ops:
                vals.append(complex(np.vdot(y0, O.to_csr() @ y0)) / nrm)
        else:
            d = req.d
            rho = y0.reshape(d, d, order="F")
            tr = np.trace(rho).real
            for O in req.e_ops:
                vals.append(complex(O.to_csr().multiply(rho.T).sum()) / tr)
        return vals

    def poll(self):
        """Advance without blocking; True when the request has finished."""
        if self.done:
            return True
        if self.kind == "se_small":
            ok = C.c_int(0)
            _lib.call("qa_event_query", self.event, C.byref(ok))
            if not ok.value:
                return False
            vals, _, info = self.small.collect(self._h)
            self._h = None
            if info[0, 0] < 0:
                self.req.error = f"integration failed (status {int(info[0, 0])})"
            else:
                v = vals[0]
                vals = v[:, :-1] / v[:, -1:].real
                first = np.asarray(self._first, dtype=np.complex128)[None, :]
                self.req.result = np.vstack([first, vals]).T
            _lib.call("qa_event_destroy", self.event)
            _lib.call("qa_stream_destroy", self.stream)
            self.done = True
            return True
        if self.kind in ("se", "me"):
            st = C.c_int(0)
            _lib.call("qa_rk_advance", self.eng.handle, C.byref(st))
            if st.value == 0:
                return False
            nf = len(self.req.e_ops) + 1
            out = np.zeros((len(self._times), nf), dtype=np.complex128)
            status = C.c_int(0)
            tr = C.c_double(0.0)
            rc = _lib.lib().qa_rk_finish(self.eng.handle, out.ctypes.data_as(C.POINTER(C.c_double)),
                                         C.byref(status), C.byref(tr))
            if rc != 0:
                self.req.error = f"integration failed (status {status.value})"
            else:
                vals = out[:, :-1] / out[:, -1:].real
                first = np.asarray(self._first, dtype=np.complex128)[None, :]
                self.req.result = np.vstack([first, vals]).T
            self.done = True
            return True
        ok = C.c_int(0)
        _lib.call("qa_event_query", self.event, C.byref(ok))
        if not ok.value:
            return False
        ex, info = self.solver.collect_async(self._mc)
        if (info[:, 0] != 0).any():
            self.req.error = f"trajectory status {sorted(set(info[:, 0].tolist()))}"
        else:
            self.req.result = ex.mean(axis=0).T
            # standard error of the ensemble mean (real parts)
            self.req.stderr = ex.real.std(axis=0).T / np.sqrt(max(ex.shape[0], 1))
        _lib.call("qa_event_destroy", self.event)
        _lib.call("qa_stream_destroy", self.stream)
        self.done = True
        return True


class SolveServer:
    """Admit by HBM budget and slot count, overlap in-flight requests on
    their own streams, retire and admit continuously."""

    def __init__(self, hbm_budget=64 << 30, max_slots=32, policy="fifo"):
        """policy: 'fifo' admits arrived requests in arrival order; 'sjf'
        admits the arrived request with the least estimated work first
        (lower median latency under overload, longer waits for the
        largest requests)."""
        if policy not in ("fifo", "sjf"):
            raise ValueError("policy must be 'fifo' or 'sjf'")
        self.policy = policy
        self.hbm_budget = int(hbm_budget)
        self.max_slots = int(max_slots)
        self.pending = []
        self.running = []
        self.finished = []
        self.in_use = 0
        self.peak_in_flight = 0

    def submit(self, req):
        self.pending.append(req)

    def _admit(self, now):
        admitted = 0
        keep = []
        order = self.pending
        if self.policy == "sjf":
            arrived = [r for r in self.pending if r.arrival <= now]
            later = [r for r in self.pending if r.arrival > now]
            order = sorted(arrived, key=lambda r: (r.work_estimate(), r.arrival)) + later
        for req in order:
            need = req.device_bytes()
            if req.arrival <= now and need > self.hbm_budget:
                # can never fit: fail it rather than wait forever
                req.error = f"needs {need} bytes, above the HBM budget {self.hbm_budget}"
                req.t_admit = req.t_done = self._clock()
                self.finished.append(req)
                continue
            if (req.arrival <= now and len(self.running) < self.max_slots
                    and self.in_use + need <= self.hbm_budget):
                req.t_admit = self._clock()
                req._bytes = need
                self.in_use += need
                try:
                    self.running.append(_Running(req))
                except Exception as e:  # a bad request fails alone
                    req.error = repr(e)
                    req.t_done = self._clock()
                    self.in_use -= need
                    self.finished.append(req)
                admitted += 1
            else:
                keep.append(req)
        self.pending = keep
        return admitted

    # queued releases are done (one device synchronisation) past these
    release_count = 4096
    release_bytes = 8 << 30

    def run(self, requests=None, realtime=True, clock=None):
        """Serve ``requests`` (arrival times in seconds from the start) until
        all have finished; returns them with result / t_done filled in.

        Device memory of retired requests is released in batches: hipFree
        waits for the whole device, so freeing each request's buffers as it
        retires would stall every request in flight."""
        for r in requests or []:
            self.submit(r)
        self.pending.sort(key=lambda r: r.arrival)
        t0 = time.perf_counter()
        self._clock = lambda: time.perf_counter() - t0
        now = lambda: (time.perf_counter() - t0) if realtime else float("inf")
        with _data.deferred_release():
            while self.pending or self.running:
                self._admit(now())
                self.peak_in_flight = max(self.peak_in_flight, len(self.running))
                still = []
                for run in self.running:
                    if run.poll():
                        run.req.t_done = time.perf_counter() - t0
                        self.in_use -= run.req._bytes
                        self.finished.append(run.req)
                    else:
                        still.append(run)
                self.running = still
                count, nbytes = _data.pending_releases()
                if count > self.release_count or nbytes > self.release_bytes:
                    _data.release_deferred()
                if not self.running and self.pending and realtime:
                    wait = min(r.arrival for r in self.pending) - now()
                    if wait > 0:
                        time.sleep(min(wait, 0.01))
        return self.finished
