"""``hip_rk``: QuTiP's explicit Runge-Kutta integrators on the GPU.

Two execution forms of the same algorithm (qutip/solver/integrator/
explicit_rk.pyx, tableaus vern7 / tsit5):

* fused (default when the right-hand side is a constant operator): the
  ``qa_rk_*`` engine of libqutip_amd runs the entire step loop, including
  step control, dense output and the output-time expectation values, on the
  device (one hipGraph replay per loop iteration);
* plain: the reference's structure -- a host loop that calls
  ``derivative(t, y)`` once per stage on device data and combines stages with
  device kernels; it accepts any derivative that works on HipDense (for
  example a QobjEvo's ``matmul_data``).

Options are those of IntegratorVern7 (qutip/solver/integrator/
qutip_integrator.py:52-60) plus ``tableau`` ('vern7' | 'tsit5'), ``fused``
(bool), ``operator_precision`` ('fp64' | 'fp32': complex64 operator
values, FP64 state, accumulation and error norm), ``operator_format``
('auto' | 'coded' | 'sell': the fused engine's operator as 2-byte codes into
value and column-offset tables when it has at most 256 of each -- the same
result bits as the plain SELL-C values and columns, a tenth of the bytes)
and ``dense_precision`` ('auto' | 'fp32' | 'bf16x6' | 'bf16x3' |
'fp64_banded': the matrix-form Lindblad RHS as MFMA products -- FP32, or
FP32 inputs split into three bf16 pieces (six products, FP32 accuracy) or
two (three products, about 16 bits: outside the 1e-6 error bar on long
solves) -- or in FP64 from the operators' diagonals; 'auto' takes the
banded form when H_nh and the collapse operators have few diagonals).
"""
import ctypes as C

import numpy as np
import scipy.sparse as sp

from .. import _lib, autotune
from .. import data as _data
from ..coefficient import DeviceProgram
from . import small as _small
from .integrator import Integrator, IntegratorException
from .tableaus import TABLEAUS

__all__ = ["IntegratorHipRK", "RKEngine", "STATUS_MESSAGES"]

AT_FRONT, INTERPOLATED, NORMAL = 2, 1, 0
TOO_MUCH_WORK, DT_UNDERFLOW, OUTSIDE_RANGE, NOT_INITIATED = -1, -2, -3, -4
STATUS_MESSAGES = {  # explicit_rk.pyx:480-496
    AT_FRONT: "Internal state at the desired time.",
    INTERPOLATED: ("Internal state past the desired time and "
                   "interpolation to step done."),
    NORMAL: "No work done.",
    TOO_MUCH_WORK: ("Too much work done in one call. Try to increase "
                    "the nsteps parameter or increasing the tolerance."),
    DT_UNDERFLOW: "Step size becomes too small. Try increasing tolerance.",
    OUTSIDE_RANGE: "Step outside available range.",
    NOT_INITIATED: "Not initialized.",
}


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


DIA_MAX_TERMS = 96  # products of diagonals the banded kernel takes (kDiaMaxTerms)
MAX_COEF = 8  # time-dependent elements of one generator on the device (kMaxCoef)
# MFMA forms of the matrix-form RHS -> qa_rk_create_dense precision codes
DENSE_PRECISIONS = {"fp32": 0, "bf16x3": 1, "bf16x6": 2}


def banded_operators(Hnh, cops):
    """The diagonals of H_nh and of each collapse operator for the banded
    matrix-form kernel (qa_rk_create_dia): (diag_op int32, diag_off int32,
    diag_vals complex128 [ndiag, N], row-indexed: A[r, r + off] at [d, r]),
    or None when the products of diagonals (2 per H_nh diagonal, the square
    of each c's count) exceed DIA_MAX_TERMS or there is nothing to multiply."""
    Hnh = np.asarray(Hnh, dtype=np.complex128)
    N = Hnh.shape[0]
    ops, offs, vals = [], [], []
    nterms = 0
    for k, A in enumerate([Hnh] + [np.asarray(c, dtype=np.complex128) for c in cops]):
        r, c = np.nonzero(A)
        od = np.unique(c - r)
        nterms += 2 * len(od) if k == 0 else len(od) ** 2
        for o in od:
            v = np.zeros(N, dtype=np.complex128)
            rr = np.arange(max(0, -o), min(N, N - o))
            v[rr] = A[rr, rr + o]
            ops.append(k)
            offs.append(int(o))
            vals.append(v)
    if not ops or nterms > DIA_MAX_TERMS:
        return None
    return (np.array(ops, dtype=np.int32), np.array(offs, dtype=np.int32),
            np.ascontiguousarray(np.array(vals)))


class RKEngine:
    """Owner of one ``qa_rk_solver`` (fused device RK for dy/dt = A y)."""

    def __init__(self, op, tableau="vern7", rtol=1e-6, atol=1e-8, nsteps=1000, first_step=0.0,
                 min_step=0.0, max_step=0.0, interpolate=True, fp32=False, lanes=None,
                 dense=None, dense_precision="fp32", coded="auto", uniform="auto", td=None):
        """``op``: the sparse generator A (dy/dt = A y), or None with
        ``dense=(H_nh, [c_ops])`` (numpy arrays) for the matrix-form
        Lindblad right-hand side: on MFMA (``dense_precision`` 'fp32',
        'bf16x6' or 'bf16x3'), or from the operators' diagonals in FP64 ('fp64_banded';
        'auto' takes it when the operators are banded, else 'fp32').  ``coded``: store A as coded
        SELL-C (2-byte codes into 256-entry value and column-offset tables,
        the same result bits) -- "auto" when A fits the tables, True to
        require it, False for the plain form; ``uniform``: the coded slices'
        layout (see data.sell_coded_arrays).  ``td``: a time-dependent
        generator A(t) = A_0 + sum_e c_e(t) A_e as (A_0, [(A_e,
        DeviceProgram of c_e), ...]) with ``op`` None: the elements go into
        one coded operator whose value table is split by element, and the
        stage kernels evaluate the programs at every stage time
        (ValueError when the stacked operator has no coded form)."""
        tab = TABLEAUS[tableau] if isinstance(tableau, str) else tableau
        self.tableau = tab
        self._a = np.ascontiguousarray(tab["a"], dtype=np.float64)
        self._b = np.ascontiguousarray(tab["b"], dtype=np.float64)
        self._c = np.ascontiguousarray(tab["c"], dtype=np.float64)
        self._e = np.ascontiguousarray(tab["e"], dtype=np.float64) if tab.get("e") is not None else None
        self._bi = np.ascontiguousarray(tab["bi"], dtype=np.float64) if tab.get("bi") is not None else None
        tab_args = (int(tab["order"]), len(self._b), len(self._c),
                    self._bi.shape[1] if self._bi is not None else 0,
                    _dp(self._a), _dp(self._b), _dp(self._c),
                    _dp(self._e) if self._e is not None else None,
                    _dp(self._bi) if self._bi is not None else None,
                    float(rtol), float(atol), int(nsteps), float(first_step), float(min_step),
                    float(max_step), int(bool(interpolate)))
        L = _lib.lib()
        h = C.c_void_p()
        if dense is not None:
            Hnh, cops = dense
            N = Hnh.shape[0]
            self.n = N * N
            self.lanes = 0
            self.sell = None
            self.coded = False
            band = None
            if dense_precision in ("auto", "fp64_banded"):
                band = banded_operators(Hnh, cops)
                if band is None and dense_precision == "fp64_banded":
                    raise ValueError("the operators are not banded enough for the banded "
                                     f"kernel (at most {DIA_MAX_TERMS} products of diagonals)")
                dense_precision = "fp64_banded" if band is not None else "fp32"
            self.dense_precision = dense_precision
        if dense is not None and band is not None:
            dop, doff, dval = band
            self._dense_host = band
            rc = L.qa_rk_create_dia(C.byref(h), N, len(cops), len(dop), _ip(dop), _ip(doff),
                                    _dp(dval.view(np.float64)), *tab_args)
            _lib.check(rc, "qa_rk_create_dia")
        elif dense is not None:
            Hf = np.asfortranarray(Hnh, dtype=np.complex128)
            cf = np.concatenate([np.asfortranarray(c, dtype=np.complex128).ravel(order="F")
                                 for c in cops]) if cops else np.zeros(1, np.complex128)
            self._dense_host = (Hf, cf)
            prec = DENSE_PRECISIONS[dense_precision]
            rc = L.qa_rk_create_dense(
                C.byref(h), N, len(cops), _dp(Hf.ravel(order="F").view(np.float64)),
                _dp(np.ascontiguousarray(cf).view(np.float64)), prec, *tab_args)
            _lib.check(rc, "qa_rk_create_dense")
        elif td is not None:
            A0, parts = td
            if not 1 <= len(parts) <= MAX_COEF:
                raise ValueError(f"1 to {MAX_COEF} time-dependent elements")
            A, elem = _data.stacked_elements([A0] + [M for M, _ in parts])
            self.n = A.shape[0]
            if lanes is None:
                lanes = autotune.lanes_for(A, False, True)
            self.lanes = lanes
            self.coded = True
            self.sell = _data.SellMatrix(A, lanes=lanes, uniform=uniform, elem=elem)
            sptr, cols, vals = self.sell.ptrs
            rc = L.qa_rk_create_coded(
                C.byref(h), self.n, lanes, self.sell.nslices, C.c_void_p(sptr),
                C.c_void_p(cols), C.c_void_p(vals), C.c_void_p(self.sell.otab_ptr),
                self.sell.vbits, *tab_args)
            _lib.check(rc, "qa_rk_create_coded")
            self.handle = h
            progs = DeviceProgram.concat([p for _, p in parts])
            start, code, consts = progs.arrays()
            vstart = np.ascontiguousarray(self.sell.vstart[1:], dtype=np.int32)
            self._td_host = (start, vstart, code, consts)
            _lib.call("qa_rk_set_coefficients", h, len(parts), _ip(start), _ip(vstart),
                      len(code), _ip(code), len(consts), _dp(consts.view(np.float64)))
        else:
            if isinstance(op, _data.HipCSR):
                A = op.to_scipy()
            else:
                A = sp.csr_matrix(op)
            if A.shape[0] != A.shape[1]:
                raise ValueError("the generator must be square")
            self.n = A.shape[0]
            if coded == "auto":
                coded = _data.sell_coded_arrays(A, 1, fp32) is not None
            coded = bool(coded)
            if lanes is None:
                lanes = autotune.lanes_for(A, fp32, coded)
            self.lanes = lanes
            self.coded = coded
            self.sell = _data.SellMatrix(A, fp32=fp32, lanes=lanes, coded=coded, uniform=uniform)
            sptr, cols, vals = self.sell.ptrs
            if coded:
                rc = L.qa_rk_create_coded(
                    C.byref(h), self.n, lanes, self.sell.nslices, C.c_void_p(sptr),
                    C.c_void_p(cols), C.c_void_p(vals), C.c_void_p(self.sell.otab_ptr),
                    self.sell.vbits, *tab_args)
                _lib.check(rc, "qa_rk_create_coded")
            else:
                rc = L.qa_rk_create(
                    C.byref(h), self.n, lanes, self.sell.nslices, C.c_void_p(sptr),
                    C.c_void_p(cols), C.c_void_p(vals), int(fp32), *tab_args)
                _lib.check(rc, "qa_rk_create")
        self.handle = h
        self.nfun = 0
        # entries between consecutive states in a states_out buffer
        self.npad = self.sell.nslices * (64 // self.lanes) if self.sell is not None else self.n
        self._y_dev = _data.HipDense((self.n, 1))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib._lib is not None:
            if not _data._defer("engine", h):
                _lib._lib.qa_rk_destroy(h)
            self.handle = None

    def set_functionals(self, kinds, fptr, rows, cols, w):
        """Linear / quadratic functionals of the state evaluated at every
        output time (qa_rk_set_functionals).  The same functionals as the
        last call are left as they are: new ones are uploaded and make the
        engine capture its loop graph again."""
        kinds = np.ascontiguousarray(kinds, dtype=np.int32)
        fptr = np.ascontiguousarray(fptr, dtype=np.int32)
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        w = np.ascontiguousarray(w, dtype=np.complex128).view(np.float64)
        key = tuple((a.shape, a.tobytes()) for a in (kinds, fptr, rows, cols, w))
        if key == getattr(self, "_fun_key", None):
            return
        self._fun_key = key
        _lib.call("qa_rk_set_functionals", self.handle, len(kinds), _ip(kinds), _ip(fptr),
                  _ip(rows), _ip(cols), _dp(w))
        self.nfun = len(kinds)

    def set_hermitian(self, on, tune=True):
        """Hermitian states on the dense engine (assume_hermitian_state):
        the upper tiles and their mirror; no effect on the other engines.
        On an MFMA engine the tile shape for that mode is then taken from
        the autotuner (measured once per size, precision and mode)."""
        _lib.call("qa_rk_set_hermitian", self.handle, int(bool(on)))
        if tune and getattr(self, "dense_precision", "fp64_banded") in ("fp32", "bf16x3",
                                                                         "bf16x6"):
            shape = autotune.dense_shape_for(self)
            if shape is not None:
                self.dense_shape = shape

    @property
    def dense_shape(self):
        """(tiled, nsplit, hsplit) of an MFMA engine."""
        t, n, h = C.c_int(), C.c_int(), C.c_int()
        _lib.call("qa_rk_get_dense_shape", self.handle, C.byref(t), C.byref(n), C.byref(h))
        return (t.value, n.value, h.value)

    @dense_shape.setter
    def dense_shape(self, shape):
        _lib.call("qa_rk_set_dense_shape", self.handle, *map(int, shape))

    @property
    def hermitian(self):
        h = C.c_int(0)
        _lib.call("qa_rk_get_hermitian", self.handle, C.byref(h))
        return bool(h.value)

    def set_state(self, t0, y0):
        y0 = y0 if isinstance(y0, _data.HipDense) else _data.HipDense.from_array(y0)
        if y0.size != self.n:
            raise ValueError(f"state has {y0.size} entries, the generator {self.n}")
        _lib.call("qa_rk_set_state", self.handle, float(t0), C.c_void_p(y0.ptr))
        self.t_set = float(t0)  # the state is the one set at t0 until a run moves it

    def run(self, times, single_step=False, states_out=None):
        """Returns (functionals [ntimes, nfun] complex, status, t_reached);
        raises IntegratorException when the integration fails."""
        times = np.ascontiguousarray(np.atleast_1d(times), dtype=np.float64)
        out = np.zeros((len(times), max(self.nfun, 1)), dtype=np.complex128)
        st = C.c_int(0)
        tr = C.c_double(0.0)
        self.t_set = None
        rc = _lib.lib().qa_rk_run(self.handle, len(times), _dp(times), int(single_step),
                                  _dp(out.view(np.float64)),
                                  C.c_void_p(states_out) if states_out else None,
                                  C.byref(st), C.byref(tr))
        if rc == _lib.QA_ERR_STATE and st.value < 0:
            raise IntegratorException(STATUS_MESSAGES.get(st.value, f"status {st.value}"))
        _lib.check(rc, "qa_rk_run")
        return out[:, :self.nfun], st.value, tr.value

    def get_state(self, copy=True):
        t = C.c_double(0.0)
        _lib.call("qa_rk_get_state", self.handle, C.byref(t), C.c_void_p(self._y_dev.ptr))
        return t.value, (self._y_dev.copy() if copy else self._y_dev)

    def stats(self):
        s = (C.c_int64 * 8)()
        _lib.call("qa_rk_stats", self.handle, s)
        keys = ["attempts", "accepted", "dense_preps", "rhs_evals", "graph_launches",
                "status", "out_idx", "phase"]
        return dict(zip(keys, [int(x) for x in s]))

    def iterate(self, count):
        self.t_set = None
        _lib.call("qa_rk_iterate", self.handle, int(count))

    def apply(self, x, out=None, t=0.0):
        """out = A(t) x (device vectors of n entries)."""
        out = _data.HipDense(x.shape) if out is None else out
        _lib.call("qa_rk_apply_at", self.handle, float(t), C.c_void_p(x.ptr), C.c_void_p(out.ptr))
        return out


class PlainRK:
    """Host-driven restatement of Explicit_RungeKutta on device data: one
    derivative call per stage, stage combines as single lincomb kernels."""

    def __init__(self, deriv, tableau, rtol, atol, nsteps, first_step, min_step, max_step,
                 interpolate):
        self.deriv = deriv
        tab = TABLEAUS[tableau] if isinstance(tableau, str) else tableau
        self.order = tab["order"]
        self.a, self.b, self.c = tab["a"], tab["b"], tab["c"]
        self.e, self.bi = tab.get("e"), tab.get("bi")
        self.rk_step, self.rk_extra = len(self.b), len(self.c)
        self.adaptive = self.e is not None
        self.interpolate = self.bi is not None and interpolate
        self.dorder = self.bi.shape[1] if self.bi is not None else 0
        self.fsal = bool(np.allclose(self.a[self.rk_step - 1, :self.rk_step], self.b, atol=1e-14))
        self.rtol, self.atol = rtol, atol
        self.max_numsteps = nsteps
        self.first_step = first_step
        self.min_step = min_step or 1e-15
        self.max_step = max_step
        self._y_prev = None
        self._status = NORMAL

    def _f(self, t, y):
        out = _data.zeros(*y.shape)
        r = self.deriv(t, y, out) if _accepts_out(self.deriv) else self.deriv(t, y)
        return out if r is None else r

    def _comb(self, base, coefs, ks, dt):
        terms = [(dt * c, k) for c, k in zip(coefs, ks) if dt * c != 0]
        out = _data.HipDense(base.shape if base is not None else ks[0].shape)
        vecs = (C.c_void_p * 16)(*[C.c_void_p(k.ptr) for _, k in terms])
        cf = (C.c_double * 16)(*[c for c, _ in terms])
        _lib.call("qa_lincomb_z", out.size, C.c_void_p(base.ptr) if base is not None else None,
                  len(terms), vecs, cf, C.c_void_p(out.ptr), None)
        return out

    def set_initial_value(self, y0, t):
        self._t = self._t_prev = self._t_front = t
        self._dt_int = 0.0
        self._y = y0
        self._y_front = y0
        self._y_prev = y0
        self.k = [None] * (self.rk_extra if self.interpolate else self.rk_step)
        if self.fsal:
            self._k_fsal = self._f(t, y0)
        self._dt_safe = self.first_step or self._estimate_first_step(t, y0)
        self._status = NORMAL

    def _estimate_first_step(self, t, y0):
        if not self.adaptive:
            return 0.0
        norm = _data.frobenius(y0)
        tol = self.atol + norm * self.rtol
        self.k[0] = self._k_fsal if self.fsal else self._f(t, y0)
        if norm <= self.atol:
            norm = 1
        tmp = _data.frobenius(self.k[0])
        fact = float(np.prod(np.arange(1, self.order + 1)))
        if tmp >= self.atol * 1e-6:
            dt1 = (tol * fact * norm ** self.order) ** (1 / (self.order + 1)) / tmp
        else:
            dt1 = (tol * fact) ** (1 / (self.order + 1)) * norm * 0.5
        yt = self._comb(y0, [1.0], [self.k[0]], dt1 / 100)
        self.k[1] = self._f(t + dt1 / 100, yt)
        tmp = _data.frobenius(self.k[1])
        dt2 = (tol * fact * norm ** self.order) ** (1 / (self.order + 1)) / tmp \
            if tmp >= self.atol * 1e-6 else dt1
        dt = min(dt1, dt2)
        if self.max_step:
            dt = min(self.max_step, dt)
        if self.min_step:
            dt = max(self.min_step, dt)
        return dt

    def integrate(self, t, step=False):
        nsteps_left = self.max_numsteps
        if self._y_prev is None:
            self._status = NOT_INITIATED
            return
        if t == self._t:
            return
        if t < self._t_prev:
            self._status = OUTSIDE_RANGE
            return
        if self.interpolate and t < self._t_front:
            if self._status != INTERPOLATED:
                self._prep_dense_out()
                self._status = INTERPOLATED
            self._y = self._interp(t)
            self._t = t
            return
        self._status = NORMAL
        if step and self._t < self._t_front and t > self._t_front:
            t = self._t_front
        while self._t_front < t and self._status >= 0:
            self._y_prev = self._y_front
            self._t_prev = self._t_front
            nsteps_left -= self._step_in_err(t, nsteps_left)
            if step:
                break
        if self._status < 0:
            return
        if self._t_front > t:
            self._prep_dense_out()
            self._status = INTERPOLATED
            self._t = t
            self._y = self._interp(t)
        else:
            self._status = AT_FRONT
            self._t = self._t_front
            self._y = self._y_front

    def _step_in_err(self, t, max_step):
        error, nsteps = 1.0, 0
        while error >= 1:
            dt = self._get_timestep(t)
            error = self._compute_step(dt)
            self._dt_int = dt
            self._recompute_safe_step(error, dt)
            if dt == self.min_step and error > 1:
                self._status = DT_UNDERFLOW
                break
            nsteps += 1
            if nsteps > max_step:
                self._status = TOO_MUCH_WORK
                break
        if self.fsal:
            self._k_fsal = self.k[self.rk_step - 1]
        return nsteps

    def _compute_step(self, dt):
        self.k[0] = self._k_fsal if self.fsal else self._f(self._t_prev, self._y_prev)
        for i in range(1, self.rk_step):
            yt = self._comb(self._y_prev, self.a[i, :i], self.k[:i], dt)
            self.k[i] = self._f(self._t_prev + self.c[i] * dt, yt)
        self._y_front = self._comb(self._y_prev, self.b, self.k[:self.rk_step], dt)
        self._t_front = self._t_prev + dt
        if not self.adaptive:
            return 0.0
        err = self._comb(None, self.e, self.k[:self.rk_step], dt)
        return _data.wrmn_error(err, self._y_front, self.atol, self.rtol)

    def _prep_dense_out(self):
        dt = self._dt_int
        for i in range(self.rk_step, self.rk_extra):
            yt = self._comb(self._y_prev, self.a[i, :i], self.k[:i], dt)
            self.k[i] = self._f(self._t_prev + self.c[i] * dt, yt)

    def _interp(self, t):
        tau = (t - self._t_prev) / self._dt_int
        bf = np.zeros(self.rk_extra)
        for i in range(self.rk_extra):
            v = 0.0
            for j in range(self.dorder - 1, -1, -1):
                v += self.bi[i, j]
                v *= tau
            bf[i] = v
        return self._comb(self._y_prev, bf, self.k[:self.rk_extra], self._dt_int)

    def _get_timestep(self, t):
        dt_needed = t - self._t_prev
        if not self.adaptive:
            return dt_needed
        if self.interpolate:
            return self._dt_safe
        if dt_needed <= self._dt_safe:
            return dt_needed
        return dt_needed / (int(dt_needed / self._dt_safe) + 1)

    def _recompute_safe_step(self, err, dt):
        if not self.adaptive:
            return
        if err == 0:
            factor = 10.0
        else:
            factor = min(10.0, max(0.2, 0.9 * err ** (-1 / (self.order + 1))))
        self._dt_safe = dt * factor
        if self.max_step:
            self._dt_safe = min(self.max_step, self._dt_safe)
        if self.min_step:
            self._dt_safe = max(self.min_step, self._dt_safe)


def _accepts_out(fn):
    return getattr(fn, "__name__", "") in ("matmul_data",)


def _as_array(op):
    """A constant operator (QobjEvo at t=0, Qobj, qutip Data or array) as a
    dense complex128 numpy array."""
    if callable(op) and not hasattr(op, "full"):
        op = op(0.0)
    if hasattr(op, "full"):
        return np.asarray(op.full(), dtype=np.complex128)
    if sp.issparse(op):
        return op.toarray().astype(np.complex128)
    if hasattr(op, "to_array"):
        return np.asarray(op.to_array(), dtype=np.complex128)
    return np.asarray(op, dtype=np.complex128)


def matrix_form_operators(owner):
    """(H_nh, [c_ops]) as dense arrays when ``owner`` is a constant
    matrix-form Lindblad right-hand side: QuTiP's LindbladMatrixForm exposes
    ``H_nh`` (QobjEvo) and ``c_ops`` (list of QobjEvo) as public attributes
    (qutip/core/cy/lindblad_matrix_form.pxd:7-8) and cannot be called
    (lindblad_matrix_form.pyx:217-222).  None for any other owner or for a
    time-dependent one."""
    if owner is None or not (hasattr(owner, "H_nh") and hasattr(owner, "c_ops")):
        return None
    Hnh, cops = owner.H_nh, list(owner.c_ops or [])
    if not all(getattr(x, "isconstant", True) for x in [Hnh] + cops):
        return None
    return _as_array(Hnh), [_as_array(c) for c in cops]


def constant_operator(derivative):
    """The constant generator behind ``derivative`` when it is the bound
    ``matmul_data`` of a constant QobjEvo (qutip/solver/integrator/_rhs.pyx:
    52-57 finds the QobjEvo the same way); None otherwise, including for
    owners that cannot be called (a matrix-form right-hand side)."""
    owner = getattr(derivative, "__self__", None)
    if owner is None or not getattr(owner, "isconstant", False):
        return None
    if hasattr(owner, "H_nh"):
        return None
    try:
        q = owner(0.0)
    except (NotImplementedError, TypeError):
        return None
    d = getattr(q, "data", None)
    if d is None:
        return None
    if isinstance(d, _data.HipCSR):
        return d.to_scipy()
    if sp.issparse(d):
        return sp.csr_matrix(d)
    if hasattr(d, "hip") and isinstance(d.hip, _data.HipCSR):  # qutip wrapper type
        return d.hip.to_scipy()
    if hasattr(d, "as_scipy"):  # qutip CSR / Dia
        return sp.csr_matrix(d.as_scipy())
    if hasattr(d, "to_array"):  # qutip Dense
        return sp.csr_matrix(d.to_array())
    return sp.csr_matrix(np.asarray(d))


def time_dependent_operator(derivative):
    """(A_0, [(A_e, DeviceProgram), ...]) when ``derivative`` is the bound
    ``matmul_data`` of a time-dependent QobjEvo of this package whose every
    coefficient has a device program (at most MAX_COEF elements); None
    otherwise (constant generators, coefficients that only run on the
    host, QuTiP's own QobjEvo whose elements are not visible from Python:
    the plain form evaluates those on the host)."""
    owner = getattr(derivative, "__self__", None)
    elements = getattr(owner, "elements", None)
    if not elements or not hasattr(owner, "qobj") or len(elements) > MAX_COEF:
        return None
    parts = []
    for q, c in elements:
        prog = c.device_program() if hasattr(c, "device_program") else None
        if prog is None:
            return None
        parts.append((sp.csr_matrix(q.to_csr()), prog))
    return sp.csr_matrix(owner.qobj.to_csr()), parts


class IntegratorHipRK(Integrator):
    """Explicit RK (vern7 / tsit5) on the GPU.  Usable with
    ``method="hip_rk"``."""
    integrator_options = {
        "atol": 1e-8,
        "rtol": 1e-6,
        "nsteps": 1000,
        "first_step": 0,
        "max_step": 0,
        "min_step": 0,
        "interpolate": True,
        "tableau": "vern7",
        "fused": True,
        "operator_precision": "fp64",
        "operator_format": "auto",
        "dense_precision": "auto",
    }
    method = "hip_rk"
    name = "hip_rk"

    def _prepare(self):
        o = self.options
        self._engine = None
        self._plain = None
        self._small = None
        self._small_start = None   # (t0, device state) of the last set_state
        self._small_final = None   # (t, device state) after a whole-solve run
        self._small_stats = None
        self._engine_pending = False  # the engine still waits for _small_start
        self._src = None    # ("op", A) or ("td", (A0, parts)): what a block engine is built from
        self._block = None  # (m, engine) for a block of m state columns
        owner = getattr(self.derivative, "__self__", None)
        dense = matrix_form_operators(owner) if o["fused"] else None
        # the matrix form's assume_hermitian_state is read at every set_state
        self._mf_owner = owner if dense is not None else None
        op = constant_operator(self.derivative) if (o["fused"] and dense is None) else None
        td = (time_dependent_operator(self.derivative)
              if (o["fused"] and dense is None and op is None
                  and o["operator_precision"] == "fp64" and o["operator_format"] != "sell")
              else None)
        if td is not None:
            try:
                self._engine = RKEngine(None, o["tableau"], o["rtol"], o["atol"], o["nsteps"],
                                        o["first_step"], o["min_step"], o["max_step"],
                                        o["interpolate"], td=td)
                self.name = "hip_rk (fused, time-dependent coded operator)"
                self._src = ("td", td)
            except ValueError:
                self._engine = None  # no coded form: the plain form below
        if self._engine is not None:
            pass
        elif dense is not None:
            self._engine = RKEngine(None, o["tableau"], o["rtol"], o["atol"], o["nsteps"],
                                    o["first_step"], o["min_step"], o["max_step"],
                                    o["interpolate"], dense=dense,
                                    dense_precision=o["dense_precision"])
            dp = self._engine.dense_precision
            self.name = ("hip_rk (fused, matrix form, FP64 banded)" if dp == "fp64_banded"
                         else f"hip_rk (fused, matrix form, {dp} MFMA)")
        elif op is not None:
            self._engine = RKEngine(op, o["tableau"], o["rtol"], o["atol"], o["nsteps"],
                                    o["first_step"], o["min_step"], o["max_step"],
                                    o["interpolate"], fp32=(o["operator_precision"] == "fp32"),
                                    coded={"auto": "auto", "coded": True,
                                           "sell": False}[o["operator_format"]])
            self.name = "hip_rk (fused, coded operator)" if self._engine.coded else "hip_rk (fused)"
            self._src = ("op", op)
            if op.shape[0] <= _small.MAX_N:
                # whole solves (Solver.run) of a small system in one workgroup
                self._small = _small.SmallSolve(
                    op, o["tableau"], o["rtol"], o["atol"], o["nsteps"], o["first_step"],
                    o["min_step"], o["max_step"], o["interpolate"],
                    fp32=(o["operator_precision"] == "fp32"))
        else:
            self._plain = PlainRK(self.derivative, o["tableau"], o["rtol"], o["atol"],
                                  o["nsteps"], o["first_step"], o["min_step"], o["max_step"],
                                  o["interpolate"])
            self.name = "hip_rk (plain)"
        # the engine of a single state column; set_state switches to a
        # block engine for several columns and back
        self._engine_base = self._engine
        self._small_base = self._small

    @property
    def fused(self):
        return self._engine is not None

    def _to_device(self, state):
        if isinstance(state, _data.HipDense):
            return state
        arr = state.full() if hasattr(state, "full") else np.asarray(state)
        return _data.HipDense.from_array(arr)

    def _block_engine(self, m):
        """The fused engine for a block of m state columns (a propagator:
        Solver.run of an operator state, propagator.py:110-200): the
        generator kron(I_m, A) acting on the column-major buffer, so the
        whole matrix is one state with one step size and one error norm,
        as the reference integrates it."""
        if self._block is not None and self._block[0] == m:
            return self._block[1]
        self._block = None
        o = self.options
        kind, src = self._src
        args = (o["tableau"], o["rtol"], o["atol"], o["nsteps"], o["first_step"], o["min_step"],
                o["max_step"], o["interpolate"])
        eye = sp.identity(m, dtype=np.complex128, format="csr")
        if kind == "op":
            eng = RKEngine(sp.kron(eye, src, format="csr"), *args,
                           fp32=(o["operator_precision"] == "fp32"),
                           coded={"auto": "auto", "coded": True,
                                  "sell": False}[o["operator_format"]])
        else:
            A0, parts = src
            eng = RKEngine(None, *args, td=(sp.kron(eye, A0, format="csr"),
                                            [(sp.kron(eye, M, format="csr"), p)
                                             for M, p in parts]))
        self._block = (m, eng)
        return eng

    def set_state(self, t, state):
        y = self._to_device(state)
        self._shape = y.shape
        base = self._engine_base
        if base is not None:
            self._engine, self._small = base, self._small_base
        if base is not None and y.size != base.n:
            n0 = base.n
            if self._src is not None and y.shape[0] == n0 and y.size % n0 == 0:
                # a block of columns (a propagator) on the fused engine
                self._engine = self._block_engine(y.size // n0)
                self._small = None
            else:
                # any other block (the matrix form): the plain form, whose
                # derivative handles matrices
                o = self.options
                self._engine = None
                self._plain = PlainRK(self.derivative, o["tableau"], o["rtol"], o["atol"],
                                      o["nsteps"], o["first_step"], o["min_step"],
                                      o["max_step"], o["interpolate"])
                self.name = "hip_rk (plain)"
        self._engine_pending = False
        if self._engine is not None and self._small is not None:
            # the whole-solve kernel starts from _small_start; the engine
            # (its first-step estimate included) is set only when a
            # step-wise call needs it
            self._engine_pending = True
        elif self._engine is not None:
            if self._mf_owner is not None:
                # MESolver._prepare_state sets it from state.isherm
                # (mesolve.py:270-276); the reference's default is True
                self._engine.set_hermitian(
                    getattr(self._mf_owner, "assume_hermitian_state", True))
            self._engine.set_state(t, y.reshape((y.size, 1)))
        else:
            self._plain.set_initial_value(y.copy(), t)
        self._small_start = (t, y.copy()) if self._engine is not None else None
        self._small_final = None
        self._is_set = True

    def _leave_small(self):
        """Continue step-wise from where a whole-solve run ended, or from
        the initial state when the engine has not been set yet."""
        if self._small_final is not None:
            t, y = self._small_final
            self._small_final = None
            self._engine.set_state(t, y.reshape((y.size, 1)))
        elif self._engine_pending:
            t, y = self._small_start
            self._engine.set_state(t, y.reshape((y.size, 1)))
        self._engine_pending = False

    def get_state(self, copy=True):
        if self._small_final is not None:
            t, y = self._small_final
            return t, (y.copy() if copy else y).reshape(self._shape)
        if self._engine_pending:
            t, y = self._small_start
            return t, (y.copy() if copy else y).reshape(self._shape)
        if self._engine is not None:
            t, y = self._engine.get_state(copy)
            return t, y.reshape(self._shape)
        y = self._plain._y
        return self._plain._t, (y.copy() if copy else y)

    def _check(self):
        st = self._plain._status
        if st < 0:
            raise IntegratorException(STATUS_MESSAGES[st])

    def integrate(self, t, copy=True):
        if self._engine is not None:
            self._leave_small()
            self._engine.run([t])
        else:
            self._plain.integrate(t, step=False)
            self._check()
        return self.get_state(copy)

    def mcstep(self, t, copy=True):
        if self._engine is not None:
            self._leave_small()
            self._engine.run([t], single_step=True)
        else:
            self._plain.integrate(t, step=True)
            self._check()
        return self.get_state(copy)

    def _wrap(self, y):
        """The Data object handed to callers (the qutip plugin wraps it)."""
        return y

    def run(self, tlist):
        """Integrator.run (qutip/solver/integrator/integrator.py:197-212):
        yields (t, state) for every t of ``tlist[1:]``.  On the fused
        engine the whole list is one device loop (qa_rk_run with states
        out): the states land in one device buffer, and each yielded state
        is a view of it that is valid until the next call of run (callers
        copy what they keep, as Result does for this method)."""
        if self._engine is None:
            for t in tlist[1:]:
                yield self.integrate(t, False)
            return
        times = np.asarray(tlist[1:], dtype=np.float64)
        if len(times) == 0:
            return
        self._leave_small()
        eng = self._engine
        step = 16 * eng.npad
        buf = _data.DeviceBuffer(step * len(times))
        self._run_states = buf  # the views below share it
        eng.run(times, states_out=buf.ptr)
        for k, t in enumerate(times):
            y = _data.HipDense((eng.n, 1), buf=buf, offset=step * k)
            yield float(t), self._wrap(y.reshape(self._shape))

    def run_functionals(self, tlist, kinds, fptr, rows, cols, w, at_start=False):
        """Fused Solver.run: functionals of the state at every time of
        ``tlist[1:]`` evaluated on the device.  Returns (values, started):
        values is [len(tlist)-1, nfun], or, when ``at_start`` is asked and
        the state was set at tlist[0] (started True), [len(tlist), nfun]
        with the functionals of the initial state at tlist[0] as row 0."""
        if self._engine is None:
            raise IntegratorException("run_functionals needs the fused engine")
        times = np.asarray(tlist[1:], dtype=np.float64)
        if self._small is not None and self._small_start is not None:
            t0, y0 = self._small_start
            started = bool(at_start) and len(tlist) > 0 and float(tlist[0]) == t0
            if started:
                times = np.asarray(tlist, dtype=np.float64)
            self._small.set_functionals(kinds, fptr, rows, cols, w)
            yf = _data.HipDense((self._small.n, 1))  # the kernel writes the final state here
            vals, _, info = self._small.run(t0, y0, times, yfin=yf)
            st = int(info[0, 0])
            if st < 0:
                raise IntegratorException(STATUS_MESSAGES.get(st, f"status {st}"))
            self._small_final = (float(times[-1]) if len(times) else t0, yf)
            self._small_stats = {"attempts": int(info[0, 3]), "dense_preps": int(info[0, 2]),
                                 "rhs_evals": self._small.rhs_evals(info[0]),
                                 "outputs": int(info[0, 1]), "status": st,
                                 "path": "whole solve in one workgroup"}
            return vals[0], started
        self._small_stats = None
        self._leave_small()
        eng = self._engine
        # an output at the set-state time is one the device loop emits first,
        # from the state itself, with the same functionals as the rest
        started = (bool(at_start) and len(tlist) > 0
                   and getattr(eng, "t_set", None) == float(tlist[0]))
        if started:
            times = np.asarray(tlist, dtype=np.float64)
        eng.set_functionals(kinds, fptr, rows, cols, w)
        vals, _, _ = eng.run(times)
        return vals, started

    def stats(self):
        if self._small_stats is not None:
            return dict(self._small_stats)
        return self._engine.stats() if self._engine is not None else {}

    @property
    def options(self):
        return self._options

    @options.setter
    def options(self, new_options):
        Integrator.options.fset(self, new_options)
        if self._options.get("tableau") not in TABLEAUS:
            raise ValueError(f"unknown tableau {self._options.get('tableau')!r}")
        if self._options.get("operator_precision") not in ("fp64", "fp32"):
            raise ValueError("operator_precision must be 'fp64' or 'fp32'")
        if self._options.get("operator_format") not in ("auto", "coded", "sell"):
            raise ValueError("operator_format must be 'auto', 'coded' or 'sell'")
        if self._options.get("dense_precision") not in ("auto", "fp64_banded", *DENSE_PRECISIONS):
            raise ValueError("dense_precision must be 'auto', 'fp32', 'bf16x6', 'bf16x3' or "
                             "'fp64_banded'")
