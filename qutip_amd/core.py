"""Host-side quantum objects: the part of QuTiP's object model the solvers
need, restated so the solvers run where qutip itself is not installed.

* ``Qobj``: matrix (scipy CSR or numpy dense, host) with ``dims``;
  arithmetic and the flags used by the solvers (qutip/core/qobj.py).
* ``QobjEvo``: constant operator wrapper with ``matmul_data`` /
  ``expect_data`` (qutip/core/cy/qobjevo.pyx:1018-1116).  Its device copy
  is made once and reused by every derivative call.
* operators and states with QuTiP's conventions (qutip/core/operators.py:
  353-505, 776; qutip/core/states.py basis): ``sigmaz = diag(1, -1)``,
  ``sigmam = [[0, 0], [1, 0]]``, ``basis(2, 0)`` spin up / excited.
* superoperators in the column-stacking convention
  (qutip/core/superoperator.py:51-229, 357-415): ``spre(A) = I (x) A``,
  ``spost(A) = A^T (x) I``, ``liouvillian``, ``lindblad_dissipator``.
"""
import numbers

import numpy as np
import scipy.sparse as sp

from . import data as _data

__all__ = ["Qobj", "QobjEvo", "LindbladMatrixForm", "basis", "ket2dm", "destroy", "create", "num", "qeye",
           "sigmax", "sigmay", "sigmaz", "sigmam", "sigmap", "tensor", "spre", "spost",
           "liouvillian", "lindblad_dissipator", "expect", "rand_herm", "rand_dm"]

HERM_ATOL = 1e-12  # settings.core['atol'] (qutip/core/options.py:183-203)


def _as_csr(m):
    return sp.csr_matrix(m, dtype=np.complex128)


class Qobj:
    __array_priority__ = 100

    def __init__(self, arr, dims=None, superrep=None, isherm=None, copy=True):
        if isinstance(arr, Qobj):
            dims = arr.dims if dims is None else dims
            arr = arr.data
        if sp.issparse(arr):
            self.data = _as_csr(arr).copy() if copy else _as_csr(arr)
        else:
            a = np.array(arr, dtype=np.complex128, copy=copy)
            if a.ndim == 1:
                a = a.reshape(-1, 1)
            self.data = a
        self.shape = self.data.shape
        if dims is None:
            dims = [[self.shape[0]], [self.shape[1]]]
        self.dims = [list(dims[0]), list(dims[1])]
        self._isherm = isherm
        self.superrep = superrep

    # -- type flags (qutip/core/dimensions.py types) ------------------------------
    @property
    def isket(self):
        return self.shape[1] == 1 and not self._is_super_dims()

    @property
    def isbra(self):
        return self.shape[0] == 1 and self.shape[1] != 1

    @property
    def isoper(self):
        return not self.isket and not self.isbra and not self.issuper

    def _is_super_dims(self):
        return isinstance(self.dims[0][0], list)

    @property
    def issuper(self):
        return self._is_super_dims() and self.shape[0] == self.shape[1]

    @property
    def isoperket(self):
        return self._is_super_dims() and self.shape[1] == 1

    @property
    def isherm(self):
        if self._isherm is None:
            if self.shape[0] != self.shape[1]:
                self._isherm = False
            else:
                d = self.full()
                self._isherm = bool(np.allclose(d, d.conj().T, atol=HERM_ATOL, rtol=0))
        return self._isherm

    # -- conversions ------------------------------------------------------------
    def full(self):
        return self.data.toarray() if sp.issparse(self.data) else np.asarray(self.data)

    def to_csr(self):
        return _as_csr(self.data)

    def dag(self):
        d = self.data.conj().T
        return Qobj(d, dims=[self.dims[1], self.dims[0]], isherm=self._isherm)

    def conj(self):
        return Qobj(self.data.conj(), dims=self.dims)

    def trans(self):
        return Qobj(self.data.T, dims=[self.dims[1], self.dims[0]])

    def tr(self):
        return complex(self.data.diagonal().sum()) if sp.issparse(self.data) else \
            complex(np.trace(self.data))

    def norm(self):
        if self.isket or self.isbra:
            return float(np.linalg.norm(self.full()))
        # trace norm for operators
        return float(np.sum(np.linalg.svd(self.full(), compute_uv=False)))

    def unit(self):
        return self * (1 / (self.norm() if (self.isket or self.isbra) else self.tr()))

    # -- arithmetic ---------------------------------------------------------------
    def __add__(self, other):
        if isinstance(other, Qobj):
            if other.shape != self.shape:
                raise ValueError("incompatible shapes")
            return Qobj(self.data + other.data, dims=self.dims)
        if isinstance(other, numbers.Number):
            if other == 0:
                return self.copy()
            return self + other * qeye_like(self)
        return NotImplemented

    __radd__ = __add__

    def __sub__(self, other):
        return self + (-1) * other

    def __rsub__(self, other):
        return (-1) * self + other

    def __neg__(self):
        return (-1) * self

    def __mul__(self, other):
        if isinstance(other, numbers.Number):
            return Qobj(self.data * other, dims=self.dims)
        if isinstance(other, Qobj):
            return self @ other
        return NotImplemented

    def __rmul__(self, other):
        if isinstance(other, numbers.Number):
            return Qobj(self.data * other, dims=self.dims)
        return NotImplemented

    def __truediv__(self, other):
        return self * (1 / other)

    def __matmul__(self, other):
        if not isinstance(other, Qobj):
            return NotImplemented
        if self.shape[1] != other.shape[0]:
            raise ValueError("incompatible shapes for matmul")
        d = self.data @ other.data
        if sp.issparse(d) and (not sp.issparse(self.data) or not sp.issparse(other.data)):
            d = d.toarray()
        return Qobj(d, dims=[self.dims[0], other.dims[1]])

    def copy(self):
        return Qobj(self.data, dims=self.dims)

    def __repr__(self):
        return f"Qobj(dims={self.dims}, shape={self.shape})"


def qeye_like(q):
    return Qobj(sp.identity(q.shape[0], dtype=np.complex128, format="csr"), dims=q.dims)


# ---------------------------------------------------------------------------
# states and operators
# ---------------------------------------------------------------------------
def basis(dims, n=0):
    d = int(np.prod(dims)) if isinstance(dims, (list, tuple)) else int(dims)
    v = np.zeros((d, 1), dtype=np.complex128)
    v[n, 0] = 1.0
    return Qobj(v, dims=[[d], [1]], copy=False)


def ket2dm(psi):
    return psi @ psi.dag()


def destroy(N, offset=0):
    d = sp.diags(np.sqrt(np.arange(offset + 1, N + offset, dtype=np.float64)), 1,
                 shape=(N, N), dtype=np.complex128, format="csr")
    return Qobj(d, dims=[[N], [N]])


def create(N, offset=0):
    return destroy(N, offset).dag()


def num(N):
    return Qobj(sp.diags(np.arange(N, dtype=np.float64), 0, format="csr", dtype=np.complex128),
                dims=[[N], [N]], isherm=True)


def qeye(N):
    return Qobj(sp.identity(N, dtype=np.complex128, format="csr"), dims=[[N], [N]], isherm=True)


def sigmax():
    return Qobj(sp.csr_matrix(np.array([[0, 1], [1, 0]], dtype=np.complex128)), isherm=True)


def sigmay():
    return Qobj(sp.csr_matrix(np.array([[0, -1j], [1j, 0]], dtype=np.complex128)), isherm=True)


def sigmaz():
    return Qobj(sp.csr_matrix(np.array([[1, 0], [0, -1]], dtype=np.complex128)), isherm=True)


def sigmam():
    return Qobj(sp.csr_matrix(np.array([[0, 0], [1, 0]], dtype=np.complex128)))


def sigmap():
    return Qobj(sp.csr_matrix(np.array([[0, 1], [0, 0]], dtype=np.complex128)))


def tensor(*args):
    if len(args) == 1 and isinstance(args[0], (list, tuple)):
        args = args[0]
    out = args[0]
    for q in args[1:]:
        if sp.issparse(out.data) or sp.issparse(q.data):
            d = sp.kron(_as_csr(out.data), _as_csr(q.data), format="csr")
        else:
            d = np.kron(out.data, q.data)
        out = Qobj(d, dims=[out.dims[0] + q.dims[0], out.dims[1] + q.dims[1]])
    return out


# ---------------------------------------------------------------------------
# superoperators (column stacking)
# ---------------------------------------------------------------------------
def _super_dims(A):
    return [[A.dims[0], A.dims[1]], [A.dims[0], A.dims[1]]]


def spre(A):
    d = A.shape[0]
    return Qobj(sp.kron(sp.identity(d, dtype=np.complex128, format="csr"), _as_csr(A.data),
                        format="csr"), dims=_super_dims(A), superrep="super")


def spost(A):
    d = A.shape[0]
    return Qobj(sp.kron(_as_csr(A.data).T, sp.identity(d, dtype=np.complex128, format="csr"),
                        format="csr"), dims=_super_dims(A), superrep="super")


def lindblad_dissipator(a, b=None):
    """D[a] rho = a rho b^dag - 1/2 {b^dag a, rho} (superoperator.py:161-229)."""
    b = a if b is None else b
    ad_b = a.dag() @ b
    return Qobj(_as_csr(spre(a).data @ spost(b.dag()).data)
                - 0.5 * _as_csr(spre(ad_b).data) - 0.5 * _as_csr(spost(ad_b).data),
                dims=_super_dims(a), superrep="super")


def liouvillian(H=None, c_ops=()):
    """-i (spre(H) - spost(H)) + sum_c D[c] (superoperator.py:51-142).  A
    time-dependent H (QobjEvo in list form) gives a QobjEvo: the constant
    Liouvillian plus -i (spre(H_e) - spost(H_e)) with each element's
    coefficient, as the reference's linear maps of QobjEvo do
    (superoperator.py:120-134)."""
    c_ops = list(c_ops or [])
    if any(isinstance(c, QobjEvo) and c.elements for c in c_ops):
        raise NotImplementedError("time-dependent collapse operators are not supported")
    c_ops = [c.qobj if isinstance(c, QobjEvo) else c for c in c_ops]
    if isinstance(H, QobjEvo):
        if H.elements:
            out = QobjEvo(liouvillian(H.qobj, c_ops))
            out.elements = [(liouvillian(q), c) for q, c in H.elements]
            return out
        H = H.qobj
    if H is not None and H.issuper:
        L = H.data.copy()
        ref = H
    elif H is not None:
        L = -1j * (_as_csr(spre(H).data) - _as_csr(spost(H).data))
        ref = H
    else:
        ref = c_ops[0]
        L = sp.csr_matrix((ref.shape[0] ** 2,) * 2, dtype=np.complex128)
    for c in c_ops:
        L = L + _as_csr(lindblad_dissipator(c).data)
    L = _as_csr(L)
    L.sum_duplicates()
    L.eliminate_zeros()
    return Qobj(L, dims=ref.dims if ref.issuper else _super_dims(ref), superrep="super")


def expect(oper, state):
    """qutip.expect for Qobj inputs (qutip/core/expect.py:79-101), on host
    data; real part when oper is Hermitian and the state is a ket or
    Hermitian."""
    O = _as_csr(oper.data)
    if state.isket:
        psi = state.full().ravel()
        val = complex(np.vdot(psi, O @ psi))
    else:
        rho = state.full()
        val = complex(O.multiply(rho.T).sum())
    if oper.isherm and (state.isket or state.isherm):
        return val.real
    return val


# ---------------------------------------------------------------------------
# random objects used by the serving workload (qutip/core/random_objects.py)
# ---------------------------------------------------------------------------
def _generator(seed):
    """A numpy Generator from a seed, or the Generator itself
    (random_objects.py:67-83)."""
    if isinstance(seed, np.random.Generator):
        return seed
    return np.random.default_rng(seed)


def rand_herm(N, density=0.3, seed=None):
    """Random Hermitian operator, the "fill" distribution of
    random_objects.py:198-290: H = (X + X^dag) / 2 with the entries of X
    uniform in [-1, 1] + i [-1, 1].  Below density 0.5 X is sparse with the
    entry count of _rand_herm_sparse (:286-308: positions drawn without
    replacement); from 0.5 up X is dense and entries are removed in
    symmetric pairs (_rand_herm_dense, :311-327).  The draws are made in
    the same order from the same numpy Generator."""
    gen = _generator(seed)
    if density < 0.5:
        target = 1 - (1 - density) ** 0.5
        num = int(max((N * N - 0.666 * N) * target + 0.666 * N * density, 1))
        data = (2 * gen.random(num) - 1) + (2 * gen.random(num) - 1) * 1j
        rows, cols = np.divmod(gen.choice(N * N, num, replace=False), N)
        M = sp.coo_matrix((data, (rows, cols)), shape=(N, N), dtype=complex).tocsr()
        M = (0.5 * (M + M.conj().T)).tocsr()
        M.sort_indices()
    else:
        M = (2 * gen.random((N, N)) - 1) + 1j * (2 * gen.random((N, N)) - 1)
        M = 0.5 * (M + M.conj().T)
        num = int(max(N * (N - 0.666) * (1 - density ** 0.5) + 0.666 * N * (1 - density), 1))
        for index in gen.choice(N * N, num, replace=False):
            r, c = divmod(int(index), N)
            M[c, r] = 0
            M[r, c] = 0
    return Qobj(M, dims=[[N], [N]], isherm=True)


def rand_dm(N, seed=None):
    """Random density matrix, the full-rank "ginibre" distribution of
    random_objects.py:508-643 (_rand_dm_ginibre): X with standard complex
    normal entries (_randnz with norm 1, :86-106: one normal(size=(N, N, 2))
    draw, real and imaginary parts last), rho = X X^dag / tr(X X^dag)."""
    gen = _generator(seed)
    g = gen.normal(size=(N, N, 2))
    X = g[..., 0] + 1j * g[..., 1]
    rho = X @ X.conj().T
    rho /= np.trace(rho)
    return Qobj(rho, dims=[[N], [N]], isherm=True)


# ---------------------------------------------------------------------------
# QobjEvo: constant operators and QuTiP's list form
# ---------------------------------------------------------------------------
class QobjEvo:
    """An operator in QuTiP's list form (qutip/core/cy/qobjevo.pyx:31-552):
    ``QobjEvo(Q)`` for a constant Qobj, or ``QobjEvo([H0, [H1, c1], ...],
    args=..., tlist=...)`` for A(t) = H0 + sum_e c_e(t) H_e, each c_e a
    string, array (with ``tlist``), function or Coefficient
    (coefficient.py:63-229).  ``qobj`` is the constant part (the sum of the
    constant elements) and ``elements`` the time-dependent ones as
    (Qobj, Coefficient) pairs (_EvoElement, _element.pyx:388-446).

    ``matmul_data(t, state)`` and ``expect_data`` take device data
    (HipDense) and run on the GPU: sum_e c_e(t) H_e @ state, the
    coefficients evaluated on the host at t (the plain form; the fused
    engine evaluates them on the device from their programs)."""

    def __init__(self, Q, args=None, tlist=None, order=3, boundary_conditions=None):
        from . import coefficient as _coef
        self.elements = []
        if isinstance(Q, QobjEvo):
            self.qobj = Q.qobj
            self.elements = [(q, c.replace_arguments(args) if args else c) for q, c in Q.elements]
        elif isinstance(Q, Qobj):
            self.qobj = Q
        elif isinstance(Q, (list, tuple)):
            const = None
            for item in Q:
                if isinstance(item, Qobj):
                    const = item if const is None else const + item
                elif isinstance(item, QobjEvo):
                    const = item.qobj if const is None else const + item.qobj
                    self.elements += item.elements
                elif isinstance(item, (list, tuple)) and len(item) == 2:
                    q, c = item
                    if not isinstance(q, Qobj):
                        raise TypeError("a list element must be [Qobj, coefficient]")
                    self.elements.append((q, _coef.coefficient(
                        c, tlist=tlist, args=args, order=order,
                        boundary_conditions=boundary_conditions)))
                else:
                    raise TypeError("QobjEvo list items are Qobj or [Qobj, coefficient]")
            if const is None:
                ref = self.elements[0][0]
                const = Qobj(sp.csr_matrix(ref.shape, dtype=np.complex128), dims=ref.dims,
                             superrep=ref.superrep)
            self.qobj = const
        else:
            raise TypeError("QobjEvo needs a Qobj or a list")
        self.dims = self.qobj.dims
        self.shape = self.qobj.shape
        for q, _ in self.elements:
            if q.shape != self.shape:
                raise ValueError("the elements of a QobjEvo must have one shape")
        self._dev = None
        self._dev_fp32 = None
        self._dev_el = None

    @property
    def isconstant(self):
        return not self.elements

    @property
    def issuper(self):
        return self.qobj.issuper

    @property
    def isoper(self):
        return self.qobj.isoper

    def __call__(self, t=0.0, args=None):
        """The operator at time t (qobjevo.pyx:437-480)."""
        out = self.qobj
        for q, c in self.elements:
            cc = c.replace_arguments(args) if args else c
            out = out + complex(cc(t)) * q
        return out

    def to_list(self):
        """The list form (qobjevo.pyx:869-890)."""
        return [self.qobj] + [[q, c] for q, c in self.elements]

    def replace_arguments(self, args=None, **kwargs):
        args = {**(args or {}), **kwargs}
        out = QobjEvo(self.qobj)
        out.elements = [(q, c.replace_arguments(args)) for q, c in self.elements]
        return out

    def arguments(self, args=None, **kwargs):
        """Replace the coefficients' arguments in place (qobjevo.pyx:
        560-590)."""
        args = {**(args or {}), **kwargs}
        if args:
            self.elements = [(q, c.replace_arguments(args)) for q, c in self.elements]

    def linear_map(self, f):
        """f applied to every element's operator, coefficients kept
        (qobjevo.pyx:818-867; f linear)."""
        out = QobjEvo(f(self.qobj))
        out.elements = [(f(q), c) for q, c in self.elements]
        out.dims, out.shape = out.qobj.dims, out.qobj.shape
        return out

    def device_data(self, fp32=False):
        if fp32:
            if self._dev_fp32 is None:
                self._dev_fp32 = _data.HipCSR(self.qobj.to_csr(), fp32=True)
            return self._dev_fp32
        if self._dev is None:
            self._dev = _data.HipCSR(self.qobj.to_csr())
        return self._dev

    def device_elements(self):
        if self._dev_el is None:
            self._dev_el = [_data.HipCSR(q.to_csr()) for q, _ in self.elements]
        return self._dev_el

    def matmul_data(self, t, state, out=None):
        """out += self(t) @ state  (qobjevo.pyx:1103-1116: the constant part,
        then every element scaled by its coefficient at t)."""
        out = _data.matmul(self.device_data(), state, 1.0, out=out)
        for (q, c), d in zip(self.elements, self.device_elements()):
            out = _data.matmul(d, state, complex(c(t)), out=out)
        return out

    def expect_data(self, t, state):
        if not self.elements:
            return _data.expect(self.device_data(), state)
        return _data.expect(_data.HipCSR(self(t).to_csr()), state)

    def _binary(self, other, sign):
        out = QobjEvo(self)
        if isinstance(other, QobjEvo):
            out.qobj = self.qobj + sign * other.qobj
            out.elements = list(self.elements) + [(sign * q, c) for q, c in other.elements]
        else:
            out.qobj = self.qobj + sign * other
        out._dev = out._dev_fp32 = out._dev_el = None
        return out

    def __add__(self, other):
        return self._binary(other, 1)

    __radd__ = __add__

    def __sub__(self, other):
        return self._binary(other, -1)

    def __mul__(self, other):
        if not isinstance(other, numbers.Number):
            return NotImplemented
        out = QobjEvo(self.qobj * other)
        out.elements = [(q * other, c) for q, c in self.elements]
        return out

    __rmul__ = __mul__

    def __neg__(self):
        return self * -1

    def dag(self):
        out = QobjEvo(self.qobj.dag())
        out.elements = [(q.dag(), c.conj()) for q, c in self.elements]
        return out

    def __matmul__(self, other):
        if isinstance(other, QobjEvo):
            if other.elements:
                raise NotImplementedError("products of two time-dependent operators")
            other = other.qobj
        out = QobjEvo(self.qobj @ other)
        out.elements = [(q @ other, c) for q, c in self.elements]
        return out

    def __rmatmul__(self, other):
        out = QobjEvo(other @ self.qobj)
        out.elements = [(other @ q, c) for q, c in self.elements]
        return out


class LindbladMatrixForm:
    """Lindblad right-hand side in matrix form (qutip/core/cy/
    lindblad_matrix_form.pyx:59-203): drho/dt = -i H_nh rho + i rho H_nh^dag
    + sum_c c rho c^dag with H_nh = H - i/2 sum_c c^dag c (:92-97).  The
    state is rho itself (N x N, not vectorised).  As in the reference
    (lindblad_matrix_form.pxd:7-10) ``H_nh`` is a QobjEvo, ``c_ops`` a list
    of QobjEvo, ``issuper`` is True and calling the object raises
    NotImplementedError (:217-222).  ``matmul_data`` runs the dense engine's
    product (MFMA, or FP64 from the diagonals of banded operators:
    ``precision``); the hip_rk integrator reads ``H_nh`` / ``c_ops`` and runs
    the whole step loop on the device."""

    issuper = True
    isoper = True

    def __init__(self, H, c_ops, precision="auto", assume_hermitian_state=True):
        if isinstance(H, QobjEvo):
            H = H.qobj
        self.c_ops = [c if isinstance(c, QobjEvo) else QobjEvo(c) for c in (c_ops or [])]
        self.dims = H.dims
        self.shape = H.shape
        self.num_collapse = len(self.c_ops)
        self.assume_hermitian_state = assume_hermitian_state
        Hnh = H.full().astype(np.complex128)
        for cop in self.c_ops:
            cm = cop.qobj.full()
            Hnh = Hnh - 0.5j * (cm.conj().T @ cm)
        self.H_nh = QobjEvo(Qobj(Hnh, dims=H.dims))
        self.precision = precision
        self._engine = None

    @property
    def isconstant(self):
        return self.H_nh.isconstant and all(c.isconstant for c in self.c_ops)

    def __call__(self, t=0.0, args=None):
        raise NotImplementedError(
            "LindbladMatrixForm cannot be called to produce an operator. "
            "Use matmul_data() to compute the RHS.")

    def matmul_data(self, t, rho, out=None):
        """out += L(rho)  (matmul_data, lindblad_matrix_form.pyx:105-203)."""
        if self._engine is None:
            from .solver.hip_rk import RKEngine, matrix_form_operators
            self._engine = RKEngine(None, dense=matrix_form_operators(self),
                                    dense_precision=self.precision)
        res = self._engine.apply(rho)
        if out is None:
            return res
        return _data.iadd(out, res, 1.0)
