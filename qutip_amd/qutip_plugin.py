"""Registration of the device path into QuTiP 5, when qutip is importable.

``register()`` wires the extension points the drop-in needs
(qutip/core/data/convert.pyx:208-404, dispatch.pyx:267-394,
qutip/solver/solver_base.py:462-492):

* data types ``HipDense`` / ``HipCSR`` as subclasses of ``qutip.core.data.Data``
  with ``to`` conversions Dense <-> HipDense and CSR <-> HipCSR, aliases
  'hip' / 'hipdense' / 'hipcsr' and the group 'hip' (dense -> HipDense,
  sparse -> HipCSR);
* dispatcher specialisations with the dispatchers' exact signatures:
  matmul, expect, inner, add, iadd, mul, imul, zeros, norm.l2, norm.frobenius,
  ode.wrmn_error, trace, column_stack, column_unstack;
* the integrator ``hip_rk`` for SESolver / MESolver / MCSolver
  (``options={'method': 'hip_rk'}``); its ``run(tlist)`` is one device loop;
* ``HipMCSolver``: qutip's MCSolver with every trajectory in one persistent
  kernel (the subclass route of nm_mcsolve.py:324-422).

qutip is not installed where this package is developed and tested, so this
module is written against the documented API and imported lazily; nothing
else in qutip_amd depends on it.
"""
import numpy as np

__all__ = ["register", "available", "HipMCSolver"]

_registered = False


def available():
    try:
        import qutip  # noqa: F401
        return True
    except Exception:  # qutip absent or broken
        return False


def register():
    """Register the device types, specialisations and the integrator."""
    global _registered
    if _registered:
        return
    import qutip
    import qutip.core.data as qd
    from qutip.solver import MCSolver, MESolver, SESolver
    from qutip.solver.integrator import Integrator as QIntegrator

    from . import data as hd
    from .solver.hip_rk import IntegratorHipRK

    class HipDenseData(qd.Data):
        """qutip Data wrapper of a HipDense buffer."""

        def __init__(self, hip, shape=None, copy=True):
            self.hip = hip.copy() if copy else hip
            super().__init__(tuple(self.hip.shape))

        @classmethod
        def sparcity(cls):
            return "dense"

        def to_array(self):
            return self.hip.to_array()

        def copy(self):
            return HipDenseData(self.hip, copy=True)

        def trace(self):
            return self.hip.trace()

        def adjoint(self):
            return HipDenseData(self.hip.adjoint(), copy=False)

        def conj(self):
            return HipDenseData(self.hip.conj(), copy=False)

        def transpose(self):
            return HipDenseData(self.hip.transpose(), copy=False)

        def __reduce__(self):
            return (_dense_from_array, (self.to_array(),))

    class HipCSRData(qd.Data):
        def __init__(self, hip, shape=None, copy=True):
            self.hip = hip.copy() if copy else hip
            super().__init__(tuple(self.hip.shape))

        @classmethod
        def sparcity(cls):
            return "sparse"

        def to_array(self):
            return self.hip.to_array()

        def copy(self):
            return HipCSRData(self.hip, copy=True)

        def trace(self):
            return self.hip.trace()

        def adjoint(self):
            return HipCSRData(self.hip.adjoint(), copy=False)

        def conj(self):
            return HipCSRData(self.hip.conj(), copy=False)

        def transpose(self):
            return HipCSRData(self.hip.transpose(), copy=False)

        def __reduce__(self):
            return (_csr_from_scipy, (self.hip.to_scipy(),))

    def dense_to_hip(m):
        return HipDenseData(hd.HipDense.from_array(m.to_array()), copy=False)

    def hip_to_dense(m):
        return qd.Dense(m.to_array())

    def csr_to_hip(m):
        return HipCSRData(hd.HipCSR(m.as_scipy()), copy=False)

    def hip_to_csr(m):
        return qd.CSR(m.hip.to_scipy())

    qd.to.add_conversions([
        (HipDenseData, qd.Dense, dense_to_hip, 1),
        (qd.Dense, HipDenseData, hip_to_dense, 1),
        (HipCSRData, qd.CSR, csr_to_hip, 1),
        (qd.CSR, HipCSRData, hip_to_csr, 1),
    ])
    qd.to.register_aliases(["hip", "hipdense", "HipDense"], HipDenseData)
    qd.to.register_aliases(["hipcsr", "HipCSR"], HipCSRData)
    qd.to.register_group(["hip"], dense=HipDenseData, sparse=HipCSRData)

    # -- specialisations: each has exactly its dispatcher's signature
    # (dispatch.pyx:287-294), e.g. matmul(left, right, scale=1)
    # (matmul.pyx:1018-1029) and zeros(rows, cols) (constant.py:18-27) ----
    def matmul_hcsr_hdense(left, right, scale=1):
        return HipDenseData(hd.matmul(left.hip, right.hip, scale), copy=False)

    # host CSR / Dia operators met on the plain path (a QobjEvo's elements
    # stay host types) are uploaded once and kept, keyed by object identity
    # (the reference's Data objects are immutable once built)
    dev_cache = {}

    def device_operator(op):
        hit = dev_cache.get(id(op))
        if hit is not None and hit[0] is op:
            return hit[1]
        csr = op if isinstance(op, qd.CSR) else qd.to(qd.CSR, op)
        dev = hd.HipCSR(csr.as_scipy())
        if len(dev_cache) >= 256:
            dev_cache.pop(next(iter(dev_cache)))
        dev_cache[id(op)] = (op, dev)
        return dev

    def matmul_host_hdense(left, right, scale=1):
        return HipDenseData(hd.matmul(device_operator(left), right.hip, scale), copy=False)

    def expect_hcsr_hdense(op, state):
        return hd.expect(op.hip, state.hip)

    def expect_host_hdense(op, state):
        return hd.expect(device_operator(op), state.hip)

    def inner_hdense(left, right, scalar_is_ket=False):
        return hd.inner(left.hip, right.hip, scalar_is_ket)

    def add_hdense(left, right, scale=1):
        return HipDenseData(hd.add(left.hip, right.hip, scale), copy=False)

    def iadd_hdense(left, right, scale=1):
        hd.iadd(left.hip, right.hip, scale)
        return left

    def mul_hdense(matrix, value):
        return HipDenseData(hd.mul(matrix.hip, value), copy=False)

    def imul_hdense(matrix, value):
        hd.imul(matrix.hip, value)
        return matrix

    def zeros_hdense(rows, cols):
        return HipDenseData(hd.zeros(rows, cols), copy=False)

    def l2_hdense(vector):
        return hd.norm_l2(vector.hip)

    def frobenius_hdense(matrix):
        return hd.frobenius(matrix.hip)

    def wrmn_hdense(diff, state, atol, rtol):
        return hd.wrmn_error(diff.hip, state.hip, atol, rtol)

    def trace_hdense(matrix):
        return matrix.hip.trace()

    def column_stack_hdense(matrix):
        return HipDenseData(hd.column_stack(matrix.hip), copy=False)

    def column_unstack_hdense(matrix, rows):
        return HipDenseData(hd.column_unstack(matrix.hip, rows), copy=False)

    qd.matmul.add_specialisations([
        (HipCSRData, HipDenseData, HipDenseData, matmul_hcsr_hdense),
        (qd.CSR, HipDenseData, HipDenseData, matmul_host_hdense),
        (qd.Dia, HipDenseData, HipDenseData, matmul_host_hdense),
    ])
    qd.expect.add_specialisations([
        (HipCSRData, HipDenseData, expect_hcsr_hdense),
        (qd.CSR, HipDenseData, expect_host_hdense),
        (qd.Dia, HipDenseData, expect_host_hdense),
    ])
    qd.inner.add_specialisations([(HipDenseData, HipDenseData, inner_hdense)])
    qd.add.add_specialisations([(HipDenseData, HipDenseData, HipDenseData, add_hdense)])
    qd.iadd.add_specialisations([(HipDenseData, HipDenseData, HipDenseData, iadd_hdense)])
    qd.mul.add_specialisations([(HipDenseData, HipDenseData, mul_hdense)])
    qd.imul.add_specialisations([(HipDenseData, HipDenseData, imul_hdense)])
    qd.zeros.add_specialisations([(HipDenseData, zeros_hdense)])
    qd.norm.l2.add_specialisations([(HipDenseData, l2_hdense)])
    qd.norm.frobenius.add_specialisations([(HipDenseData, frobenius_hdense)])
    qd.ode.wrmn_error.add_specialisations([(HipDenseData, HipDenseData, wrmn_hdense)])
    qd.trace.add_specialisations([(HipDenseData, trace_hdense)])
    qd.column_stack.add_specialisations([(HipDenseData, HipDenseData, column_stack_hdense)])
    qd.column_unstack.add_specialisations([(HipDenseData, HipDenseData, column_unstack_hdense)])

    # -- integrator -------------------------------------------------------------
    class QutipIntegratorHipRK(IntegratorHipRK, QIntegrator):
        """hip_rk as a subclass of qutip's own Integrator."""

        def _to_device(self, state):
            if isinstance(state, HipDenseData):
                return state.hip
            return hd.HipDense.from_array(state.to_array())

        def get_state(self, copy=True):
            t, y = IntegratorHipRK.get_state(self, copy)
            return t, HipDenseData(y, copy=False)

        def _wrap(self, y):
            return HipDenseData(y, copy=False)

    for cls in (SESolver, MESolver, MCSolver):
        cls.add_integrator(QutipIntegratorHipRK, "hip_rk")

    global HipMCSolver
    HipMCSolver = _make_mc_solver(MCSolver)
    _registered = True


HipMCSolver = None  # set by register(): qutip's MCSolver with batched trajectories


def _mirror(q):
    """A qutip Qobj (or constant QobjEvo) as this package's host Qobj."""
    import scipy.sparse as sp

    from . import core
    if hasattr(q, "isconstant") and not hasattr(q, "full"):
        if not q.isconstant:
            raise NotImplementedError("the batched trajectory kernel needs constant operators")
        q = q(0)
    try:
        m = sp.csr_matrix(q.to("CSR").data.as_scipy())
    except Exception:  # a data type without a CSR conversion
        m = np.asarray(q.full())
    return core.Qobj(m, dims=q.dims)


def _make_mc_solver(MCSolver):
    """A subclass of qutip's MCSolver (qutip/solver/mcsolve.py:424-821) whose
    ``run`` advances all trajectories in the persistent trajectory kernel
    (qa_mc_run) instead of one MCIntegrator per trajectory: the route the
    reference leaves for it is a subclass overriding ``run``, as
    NonMarkovianMCSolver does (nm_mcsolve.py:324-422).  Options are qutip's
    plus the device keys of qutip_amd's MCSolver ('precision', 'max_draws',
    'max_jumps', 'grid'); the result is qutip_amd's McResult (average_expect,
    std_expect, runs_expect, col_times, col_which, seeds)."""
    from .solver.mcsolve import MCSolver as BatchMC
    from .solver.tableaus import TABLEAUS

    class HipMCSolver(MCSolver):
        name = "mcsolve (hip batched)"

        def __init__(self, H, c_ops, *, options=None):
            opts = dict(options or {})
            dev = {k: opts.pop(k) for k in ("precision", "max_draws", "max_jumps", "grid")
                   if k in opts}
            super().__init__(H, c_ops, options=opts)
            mine = {k: v for k, v in self.options.items() if k in BatchMC.solver_options}
            if mine.get("method") not in TABLEAUS:
                mine["method"] = "vern7"
            mine.update(dev)
            self._batch = BatchMC(_mirror(H), [_mirror(c) for c in c_ops], options=mine)

        def run(self, state, tlist, ntraj=1, *, args=None, e_ops=(), target_tol=None,
                timeout=None, seeds=None):
            if isinstance(e_ops, dict):
                e_ops = list(e_ops.values())
            elif e_ops is None:
                e_ops = []
            elif not isinstance(e_ops, (list, tuple)):
                e_ops = [e_ops]
            if any(callable(O) and not hasattr(O, "dims") for O in e_ops):
                raise NotImplementedError("the batched trajectory kernel takes operator e_ops")
            return self._batch.run(_mirror(state), tlist, ntraj,
                                   e_ops=[_mirror(O) for O in e_ops], seeds=seeds,
                                   target_tol=target_tol)

    return HipMCSolver


def _dense_from_array(arr):
    import qutip.core.data as qd
    return qd.to("hip", qd.Dense(np.asarray(arr)))


def _csr_from_scipy(m):
    import qutip.core.data as qd
    return qd.to("hipcsr", qd.CSR(m))
