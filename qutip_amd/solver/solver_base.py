"""Solver base class: state preparation, the output loop and the integrator
registry (qutip/solver/solver_base.py:24-492), running on device data.

Two output paths, both giving the reference's Result semantics:

* fused -- when the integrator is the fused ``hip_rk`` engine and every
  e_op is an operator, the expectation values (and the normalisation
  tr(rho) or ||psi||^2) are device "functionals" evaluated inside the RK
  engine's output step;
* generic -- the reference's loop over ``integrator.run(tlist)``, with the
  expectation values computed by the device ``expect`` kernels on each
  returned state.

Normalisation follows _prepare_state / _restore_state
(solver_base.py:81-157): when ``normalize_output`` is set and the initial
state is normalised (|norm - 1| <= 1e-12), states are divided by tr(rho)
(density matrices) or ||psi|| (kets).  Real casting follows
qutip/core/expect.py:95-100: the real part is returned for a Hermitian e_op
when the state is a ket or a density matrix evolved by a
Hermiticity-preserving generator (mesolve).
"""
import time

import numpy as np
import scipy.sparse as sp

from .. import data as _data
from ..core import Qobj, QobjEvo
from .integrator import Integrator
from .result import Result

__all__ = ["Solver"]

NORM_ATOL = 1e-12  # settings.core['atol']


class Solver:
    name = ""
    _avail_integrators = {}
    _vectorize_state = True
    _preserves_hermiticity = False
    solver_options = {
        "progress_bar": "",
        "progress_kwargs": {},
        "store_final_state": False,
        "store_states": None,
        "normalize_output": True,
        "method": "hip_rk",
    }

    def __init__(self, rhs, *, options=None):
        if isinstance(rhs, Qobj):
            rhs = QobjEvo(rhs)
        if not hasattr(rhs, "matmul_data"):
            raise TypeError("The rhs must be a Qobj, QobjEvo or LindbladMatrixForm")
        self.rhs = rhs
        self._options = {}
        self.options = options
        t0 = time.time()
        self._integrator = self._get_integrator()
        self._init_time = time.time() - t0
        self.stats = self._initialize_stats()

    # -- options (solver_base.py:341-454): unknown keys raise KeyError ---------
    @property
    def options(self):
        return self._options

    @options.setter
    def options(self, new):
        new = dict(new or {})
        method = new.get("method", self._options.get("method", self.solver_options["method"]))
        integ = self.avail_integrators().get(method)
        if integ is None:
            raise ValueError(f"Integrator method {method!r} not supported by {self.name}")
        allowed = set(self.solver_options) | set(integ.integrator_options)
        extra = set(new) - allowed
        if extra:
            raise KeyError(f"Options {sorted(extra)} are not supported.")
        merged = {**self.solver_options, **integ.integrator_options, **self._options, **new}
        changed_method = merged.get("method") != self._options.get("method")
        self._options = merged
        if getattr(self, "_integrator", None) is not None:
            if changed_method:
                self._integrator = self._get_integrator()
            else:
                self._integrator.options = merged
                self._integrator.reset(hard=True)

    @classmethod
    def avail_integrators(cls):
        if cls is Solver:
            return cls._avail_integrators.copy()
        out = {}
        for base in reversed(cls.__mro__):
            out.update(getattr(base, "_avail_integrators", {}))
        return out

    @classmethod
    def add_integrator(cls, integrator, key):
        """Register an Integrator subclass under ``method=key``
        (solver_base.py:462-492)."""
        if not issubclass(integrator, Integrator):
            raise TypeError(f"The integrator must be a subclass of Integrator")
        cls._avail_integrators = {**cls.__dict__.get("_avail_integrators", {}), key: integrator}

    def _get_integrator(self):
        method = self._options["method"]
        integ = self.avail_integrators()[method]
        if integ.rhs_format == "matrix":
            return integ(self.rhs(0).data, self._options)
        return integ(self.rhs.matmul_data, self._options)

    def _initialize_stats(self):
        return {"method": self._integrator.name, "init time": self._init_time,
                "preparation time": 0.0, "run time": 0.0}

    # -- state preparation (solver_base.py:81-157) ------------------------------
    def _prepare_state(self, state):
        if (self.rhs.issuper or getattr(self, "_matrix_form", False)) and state.isket:
            # ket2dm(psi) = |psi><psi| as one outer product: the same
            # entries, without the operator product of Qobj
            psi = np.asarray(state.full()).ravel()
            arr = np.outer(psi, psi.conj())
            self._state_is_ket = False
            self._state_dims = [list(state.dims[0]), list(state.dims[0])]
            self._state_shape = arr.shape
            self._normalize_output = bool(
                self._options.get("normalize_output", False)
                and abs(float(np.vdot(psi, psi).real) - 1) <= NORM_ATOL)
            self._vectorized = bool(self._vectorize_state and self.rhs.issuper)
            if self._vectorized:
                return _data.HipDense.from_array(arr.ravel(order="F").reshape(-1, 1))
            return _data.HipDense.from_array(arr)
        if state.isket:
            norm = state.norm()
        elif state.shape[0] == state.shape[1]:
            norm = state.tr()
        else:
            norm = -1
        self._state_is_ket = state.isket
        self._state_dims = state.dims
        self._state_shape = state.shape
        self._normalize_output = bool(
            self._options.get("normalize_output", False)
            and abs(norm - 1) <= NORM_ATOL)
        arr = state.full()
        # an operator evolves as vec(rho) under a superoperator generator; a
        # superoperator state (a propagator) stays a matrix (solver_base.py:
        # 87-95)
        self._vectorized = bool(self._vectorize_state and self.rhs.issuper and not state.issuper)
        if self._vectorized:
            vec = arr.ravel(order="F").reshape(-1, 1)
            return _data.HipDense.from_array(vec)
        return _data.HipDense.from_array(arr)

    def _restore_state(self, data, copy=True):
        arr = data.to_array()
        if self._vectorized:
            arr = arr.reshape(self._state_shape, order="F")
        arr = arr.reshape(self._state_shape)
        q = Qobj(arr, dims=self._state_dims, copy=False)
        if self._normalize_output:
            q = q * (1 / (q.norm() if q.isket else q.tr()))
        return q

    # -- e_ops ----------------------------------------------------------------
    def _eop_list(self, e_ops):
        if e_ops is None:
            return []
        if isinstance(e_ops, dict):
            return list(e_ops.values())
        if isinstance(e_ops, (Qobj, QobjEvo)) or callable(e_ops):
            return [e_ops]
        return list(e_ops)

    def _functionals(self, ops):
        """Device functionals for operator e_ops plus the norm functional
        last; the arrays of the last e_ops list are kept (Qobj data is
        treated as immutable, as its cached isherm already is)."""
        key = (self._state_is_ket, self._state_shape,
               tuple((id(O), id(O.data)) for O in ops))
        hit = getattr(self, "_fun_cache", None)
        if hit is not None and hit[0] == key:
            return hit[2]
        f = self._build_functionals(ops)
        self._fun_cache = (key, [(O, O.data) for O in ops], f)  # refs keep the ids valid
        return f

    def _build_functionals(self, ops):
        kinds, fptr, rows, cols, ws = [], [0], [], [], []
        if self._state_is_ket:
            n = self._state_shape[0]
            for O in ops:
                A = sp.coo_matrix(O.to_csr())
                kinds.append(1)
                rows.append(A.row)
                cols.append(A.col)
                ws.append(A.data)
                fptr.append(fptr[-1] + A.nnz)
            kinds.append(1)
            rows.append(np.arange(n))
            cols.append(np.arange(n))
            ws.append(np.ones(n, dtype=np.complex128))
            fptr.append(fptr[-1] + n)
        else:
            d = self._state_shape[0]
            for O in ops:
                A = sp.coo_matrix(O.to_csr())
                kinds.append(0)
                rows.append(np.zeros(A.nnz, dtype=np.int64))
                cols.append(A.row * d + A.col)  # rho[c, r] sits at r * d + c
                ws.append(A.data)
                fptr.append(fptr[-1] + A.nnz)
            kinds.append(0)
            rows.append(np.zeros(d, dtype=np.int64))
            cols.append(np.arange(d) * (d + 1))
            ws.append(np.ones(d, dtype=np.complex128))
            fptr.append(fptr[-1] + d)
        cat = lambda xs, dt: np.concatenate(xs).astype(dt) if xs else np.zeros(0, dt)
        return (np.asarray(kinds, np.int32), np.asarray(fptr, np.int32),
                cat(rows, np.int32), cat(cols, np.int32), cat(ws, np.complex128))

    def _cast(self, O, val):
        herm_state = self._state_is_ket or self._preserves_hermiticity
        if O.isherm and herm_state:
            return val.real
        return val

    def _values_from_functionals(self, ops, vals):
        """One array per e_op over the times of ``vals`` (_cast applied to
        the whole array)."""
        norm = vals[:, len(ops)]
        out = []
        for j, O in enumerate(ops):
            v = vals[:, j]
            if self._normalize_output:
                v = v / norm.real
            out.append(self._cast(O, v))
        return out

    @staticmethod
    def _check_eops(ops, d):
        """Operator e_ops must act on the state's space (state dims[0] = d):
        the reference's expect raises ValueError on a dims mismatch
        (qutip/core/expect.py:85-90), and the device functionals index the
        state by the operator's rows and columns."""
        for O in ops:
            if isinstance(O, (Qobj, QobjEvo)):
                if list(O.dims[1]) != list(d) or O.shape[0] != O.shape[1]:
                    raise ValueError(f"incompatible dimensions {O.dims[1]} and {d}")
            elif not callable(O):
                raise TypeError(f"{O!r} has unsupported type {type(O)!r}.")

    # -- run ------------------------------------------------------------------
    def run(self, state0, tlist, *, e_ops=None, args=None):
        t_start = time.time()
        self._argument(args)
        tlist = np.asarray(tlist, dtype=np.float64)
        ops = self._eop_list(e_ops)
        self._check_eops(ops, state0.dims[0])
        y0 = self._prepare_state(state0)
        integ = self._integrator
        integ.set_state(tlist[0], y0)
        stats = self._initialize_stats()
        res = Result(e_ops, self._options, solver=self.name, stats=stats)
        store_states = bool(self._options.get("store_states"))
        if self._options.get("store_states") is None:
            store_states = len(ops) == 0
        store_final = bool(self._options.get("store_final_state")) or store_states
        stats["preparation time"] = time.time() - t_start
        t_run = time.time()
        fused = (getattr(integ, "fused", False) and all(isinstance(O, Qobj) for O in ops)
                 and not store_states)
        if fused:
            # the device loop; the initial point comes from the same
            # functionals when the whole-solve kernel evaluates them at
            # tlist[0], else from the host-prepared state
            vals, started = None, False
            if len(tlist) > 1:
                vals, started = integ.run_functionals(tlist, *self._functionals(ops),
                                                      at_start=True)
            if started:
                res.add_expect_block(tlist, self._values_from_functionals(ops, vals))
            else:
                res.add_expect_values(tlist[0], self._expect_on_device(ops, y0, tlist[0]))
                if vals is not None:
                    res.add_expect_block(tlist[1:], self._values_from_functionals(ops, vals))
            if store_final:
                res.final_state = self._restore_state(integ.get_state(copy=False)[1])
            stats.update(integ.stats())
        else:
            _, st = integ.get_state(copy=False)
            self._add_generic(res, tlist[0], st, ops, store_states, store_final)
            for t, st in integ.run(tlist):
                self._add_generic(res, t, st, ops, store_states, store_final)
        stats["run time"] = time.time() - t_run
        return res

    def _state_view(self, st):
        if self._vectorized:
            d = self._state_shape[0]
            return st.reshape((d, d))
        return st

    def _expect_on_device(self, ops, st, t):
        """e_ops values of the state at time t: operators by the device
        expect kernels, QobjEvo at O(t), callables as f(t, state)
        (result.py:247-262)."""
        view = self._state_view(st)
        vals = []
        if self._normalize_output:
            if self._state_is_ket:
                nrm = _data.frobenius(view) ** 2
            else:
                nrm = view.trace().real
        for O in ops:
            if callable(O) and not isinstance(O, (Qobj, QobjEvo)):
                vals.append(O(float(t), self._restore_state(st)))
                continue
            Oq = O(float(t)) if isinstance(O, QobjEvo) else O
            v = _data.expect(_data.HipCSR(Oq.to_csr()), view)
            if self._normalize_output:
                v = v / nrm
            vals.append(self._cast(Oq, v))
        return vals

    def _add_generic(self, res, t, st, ops, store_states, store_final):
        res.add_expect_values(t, self._expect_on_device(ops, st, t))
        if store_states or store_final:
            res.add_state(self._restore_state(st), store_states, store_final)

    # -- step interface (solver_base.py:228-280) -------------------------------
    def start(self, state0, t0):
        self._integrator.set_state(t0, self._prepare_state(state0))

    def _argument(self, args):
        """Replace the coefficients' arguments of the rhs and rebuild the
        integrator's device programs (solver_base.py:456-460)."""
        if args and getattr(self.rhs, "elements", None):
            self.rhs.arguments(args)
            self._integrator = self._get_integrator()

    def step(self, t, *, args=None, copy=True):
        _, state = self._integrator.integrate(t, copy=False)
        return self._restore_state(state, copy=copy)
