"""Whole solves of small systems on the device.

For a constant generator with n <= 1024 the fused engine of ``hip_rk.py``
is launch-bound: every RK iteration is about twenty kernel launches for a
few hundred state entries.  ``SmallSolve`` runs the entire integration of
one system in one workgroup instead (``qa_rk_small_run``): the stages, the
error norm and step control, dense output at the output times
(qutip/solver/integrator/explicit_rk.pyx:278-331) and the functionals that
give the e_ops values (result.py:303-340).  Many independent systems (initial
states, and optionally operators of the same sparsity width) run in one
launch, one workgroup each.
"""
import ctypes as C

import numpy as np
import scipy.sparse as sp

from .. import _lib
from .. import data as _data
from .mcsolve import _ell_host
from .tableaus import TABLEAUS

__all__ = ["SmallSolve", "MAX_N"]

MAX_N = 1024

_vp = lambda p: C.c_void_p(p)


def _pack_tableau(name, interpolate):
    tab = TABLEAUS[name]
    size = _lib.lib().qa_mc_tableau_size()
    buf = np.zeros(size, dtype=np.uint8)
    a, b, c, e, bi = (np.ascontiguousarray(tab[k], dtype=np.float64)
                      for k in ("a", "b", "c", "e", "bi"))
    dp = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))
    _lib.call("qa_mc_pack_tableau", buf.ctypes.data_as(C.c_void_p), int(tab["order"]), len(b),
              len(c), bi.shape[1], dp(a), dp(b), dp(c), dp(e), dp(bi), int(bool(interpolate)))
    return _data.DeviceBuffer.from_array(buf), tab


class SmallSolve:
    """dy/dt = A y for one or many systems of size n <= 1024, each integrated
    by one workgroup.  ``A`` is one operator, or a list of operators (one per
    system) of the same size."""

    def __init__(self, A, tableau="vern7", rtol=1e-6, atol=1e-8, nsteps=1000, first_step=0.0,
                 min_step=0.0, max_step=0.0, interpolate=True, fp32=False):
        ops = A if isinstance(A, (list, tuple)) else [A]
        ops = [sp.csr_matrix(O, dtype=np.complex128) for O in ops]
        n = ops[0].shape[0]
        if any(O.shape != (n, n) for O in ops):
            raise ValueError("every generator must be square and of the same size")
        if n > MAX_N:
            raise ValueError(f"SmallSolve handles n <= {MAX_N} (got {n})")
        self.n = n
        self.fp32 = bool(fp32)
        W = max(-(-int(np.diff(O.indptr).max()) // 4) * 4 for O in ops)
        cols, vals = [], []
        for O in ops:
            w, col, val = _ell_host(O, n, self.fp32)
            if w < W:  # pad to the common width (value 0, own row as column)
                col = np.vstack([col[:max(w, 1)] if w else col[:0],
                                 np.tile(np.arange(n, dtype=np.int32), (W - w, 1))])[:W]
                val = np.vstack([val[:w], np.zeros((W - w, n), dtype=val.dtype)])
            cols.append(col[:max(W, 1)])
            vals.append(val[:max(W, 1)])
        self.width = W
        self.nops = len(ops)
        self.stride = max(W, 1) * n if self.nops > 1 else 0
        self._col = _data.DeviceBuffer.from_array(np.ascontiguousarray(np.concatenate(cols)))
        self._val = _data.DeviceBuffer.from_array(np.ascontiguousarray(np.concatenate(vals)))
        self._tab, self.tableau = _pack_tableau(tableau, interpolate)
        self.opts = dict(rtol=float(rtol), atol=float(atol), nsteps=int(nsteps),
                         first_step=float(first_step), min_step=float(min_step),
                         max_step=float(max_step))
        self.set_functionals(np.zeros(0, np.int32), np.zeros(1, np.int32), np.zeros(0, np.int32),
                             np.zeros(0, np.int32), np.zeros(0, np.complex128))

    def set_functionals(self, kinds, fptr, rows, cols, w):
        """Linear (kind 0: sum w y[col]) or quadratic (kind 1:
        sum w conj(y[row]) y[col]) functionals, CSR-like by ``fptr``."""
        arrs = [np.ascontiguousarray(x, dtype=dt) for x, dt in
                ((kinds, np.int32), (fptr, np.int32), (rows, np.int32), (cols, np.int32),
                 (w, np.complex128))]
        key = tuple((a.shape, a.tobytes()) for a in arrs)
        if key == getattr(self, "_fun_key", None):
            return  # the same functionals are on the device already
        # the kernel reads the functionals from the device without bounds
        # checks, so their host copies are checked against n first
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
        _lib.host_call("qa_functionals_check", self.n, len(arrs[0]), ip(arrs[0]), ip(arrs[1]),
                       ip(arrs[2]), ip(arrs[3]))
        self._fun_key = key
        self.nfun = len(kinds)
        up = lambda x, dt: _data.DeviceBuffer.from_array(
            np.ascontiguousarray(x if len(x) else np.zeros(1), dtype=dt))
        self._fk = up(kinds, np.int32)
        self._fp = up(fptr, np.int32)
        self._fr = up(rows, np.int32)
        self._fc = up(cols, np.int32)
        self._fw = up(w, np.complex128)

    def _scratch(self, name, nbytes, reuse):
        """A device buffer of at least nbytes: a fresh one, or (reuse, for the
        synchronous run()) the one kept from the previous call when it is big
        enough -- allocations and frees synchronise the device."""
        if not reuse:
            return _data.DeviceBuffer(nbytes)
        bufs = self.__dict__.setdefault("_bufs", {})
        b = bufs.get(name)
        if b is None or b.nbytes < nbytes:
            b = bufs[name] = _data.DeviceBuffer(nbytes)
        return b

    def launch(self, t0, y0, times, grid=0, stream=None, reuse=False, yfin=None):
        """Enqueue the solves on ``stream`` and return a handle that keeps the
        device buffers alive; ``collect`` reads the results once the stream
        has reached this point.  ``reuse``: inputs and outputs go to buffers
        kept on this object (only for a launch that is collected before the
        next one, as run() does).  ``yfin``: a device buffer (``.ptr``) of at
        least nsys * n complex128 for the final states, else a new one."""
        times = np.ascontiguousarray(np.atleast_1d(times), dtype=np.float64)
        nt = len(times)
        if isinstance(y0, _data.HipDense):
            nsys = y0.size // self.n
            y0buf = y0
        else:
            y = np.ascontiguousarray(np.asarray(y0, dtype=np.complex128).reshape(-1, self.n))
            nsys = y.shape[0]
            y0buf = self._scratch("y0", y.nbytes, reuse)
            y0buf.upload(y)
        if self.nops > 1 and nsys > self.nops:
            raise ValueError(f"{self.nops} operators for {nsys} systems")
        tb = self._scratch("times", 8 * max(nt, 1), reuse)
        if nt:
            tb.upload(times)
        H = {"nsys": nsys, "nt": nt, "stream": stream, "y0": y0buf, "times": tb,
             "out": self._scratch("out", 16 * max(nsys * nt * self.nfun, 1), reuse),
             "yfin": yfin if yfin is not None else _data.DeviceBuffer(16 * nsys * self.n),
             "info": self._scratch("info", 16 * max(nsys, 1), reuse)}
        o = self.opts
        rc = _lib.lib().qa_rk_small_run(
            self.n, nsys, float(t0), nt, _vp(H["times"].ptr), _vp(y0buf.ptr), self.width,
            _vp(self._col.ptr), _vp(self._val.ptr), self.stride, self.nfun, _vp(self._fk.ptr),
            _vp(self._fp.ptr), _vp(self._fr.ptr), _vp(self._fc.ptr), _vp(self._fw.ptr),
            _vp(self._tab.ptr), o["rtol"], o["atol"], o["nsteps"], o["first_step"],
            o["min_step"], o["max_step"], int(self.fp32), _vp(H["out"].ptr),
            _vp(H["yfin"].ptr), _vp(H["info"].ptr), int(grid), stream)
        _lib.check(rc, "qa_rk_small_run")
        return H

    def collect(self, H):
        """(functionals [nsys, nt, nfun], final states DeviceBuffer [nsys, n],
        info [nsys, 4]) of a finished ``launch``."""
        if H["stream"] is not None:
            _lib.call("qa_stream_sync", H["stream"])
        vals = np.zeros((H["nsys"], H["nt"], self.nfun), dtype=np.complex128)
        if vals.size:
            H["out"].download(vals)
        inf = np.zeros((H["nsys"], 4), dtype=np.int32)
        H["info"].download(inf)
        return vals, H["yfin"], inf

    def run(self, t0, y0, times, grid=0, stream=None, yfin=None):
        """Integrate every system from (t0, y0[s]) through ``times`` (system
        s uses operator s when there is one per system).  Returns
        (functionals [nsys, len(times), nfun] complex, final states as a
        DeviceBuffer [nsys, n] complex128, info [nsys, 4]: status, outputs
        written, dense-output preparations, attempts)."""
        return self.collect(self.launch(t0, y0, times, grid, stream, reuse=True, yfin=yfin))

    def rhs_evals(self, info_row):
        """RHS evaluations of one system from its info row (the reference's
        count: two at the start (the first-step estimate, or the FSAL slot
        and one estimate product), the step stages of every attempt and the
        extra stages of every dense-output preparation)."""
        tab = self.tableau
        a, b = np.asarray(tab["a"]), np.asarray(tab["b"])
        rk_step, rk_extra = len(b), len(tab["c"])
        fsal = bool(np.allclose(a[rk_step - 1, :rk_step], b))
        per_attempt = rk_step - 1 if fsal else rk_step
        return int(2 + per_attempt * info_row[3] + (rk_extra - rk_step) * info_row[2])
