"""Device data types and their data-layer operations.

``HipDense`` and ``HipCSR`` mirror QuTiP's ``Dense`` and ``CSR`` data types
(qutip/core/data/dense.pxd:9-33, csr.pxd:17-30) with their storage in HBM:

* HipDense: complex128, column-major ("Fortran" order), so a density matrix
  buffer *is* its column-stacked vector (qutip/core/superoperator.py:291-302)
  and stack/unstack are free reshapes;
* HipCSR: int32 row pointers / columns (QuTiP's ``idxint``), complex128 or
  complex64 values; a SELL-64 copy is built on first use by the fused RK
  engine.

The module-level functions carry the signatures of the reference
dispatchers they specialise (``matmul(left, right, scale=1)``,
``expect(op, state)``, ``inner(left, right, scalar_is_ket=False)``, ...);
qutip_amd.qutip_plugin registers them into qutip.core.data when qutip is
importable.  All arithmetic runs in the gfx950 kernels; the host only moves
data in and out (``from_array`` / ``to_array``).
"""
import ctypes as C
import numbers

import numpy as np
import scipy.sparse as sp

from . import _lib

__all__ = ["HipDense", "HipCSR", "DeviceBuffer", "deferred_release", "release_deferred",
           "matmul", "add", "iadd", "mul", "imul",
           "expect", "inner", "norm_l2", "frobenius", "wrmn_error", "zeros", "trace",
           "column_stack", "column_unstack", "to_hip", "to_numpy"]


# Deferred release.  hipFree waits for the whole device, so a host loop that
# frees while other work is in flight (the solve server) serialises on every
# release.  Inside ``deferred_release()`` the buffers and engines that die
# are queued instead, and ``release_deferred()`` frees the queue after one
# device synchronisation.
_deferred = None
_deferred_bytes = 0


def _defer(kind, obj, nbytes=0):
    """Queue a release when deferral is on; True when queued."""
    global _deferred_bytes
    if _deferred is None:
        return False
    _deferred.append((kind, obj))
    _deferred_bytes += nbytes
    return True


class deferred_release:
    """Context manager: releases inside it are queued, then done at exit
    (or earlier, by ``release_deferred()``)."""

    def __enter__(self):
        global _deferred
        if _deferred is None:
            _deferred = []
            self._owner = True
        else:
            self._owner = False
        return self

    def __exit__(self, *exc):
        global _deferred
        if self._owner:
            release_deferred()
            _deferred = None
        return False


def pending_releases():
    """(queued releases, queued buffer bytes)."""
    return (0 if _deferred is None else len(_deferred)), _deferred_bytes


def release_deferred():
    """Synchronise the device once and free everything queued."""
    global _deferred_bytes
    if not _deferred or _lib._lib is None:
        return
    L = _lib._lib
    L.qa_device_sync()
    _deferred_bytes = 0
    while _deferred:
        kind, obj = _deferred.pop()
        if kind == "buffer":
            L.qa_free(C.c_void_p(obj))
        else:
            L.qa_rk_destroy(obj)


class DeviceBuffer:
    """A block of HBM owned by this object (freed on garbage collection, or
    queued for release inside ``deferred_release()``)."""
    __slots__ = ("ptr", "nbytes", "__weakref__")

    def __init__(self, nbytes):
        L = _lib.lib()
        p = C.c_void_p()
        _lib.check(L.qa_malloc(C.byref(p), C.c_size_t(max(int(nbytes), 1))), "qa_malloc")
        self.ptr = p.value
        self.nbytes = int(nbytes)

    def __del__(self):
        p = getattr(self, "ptr", None)
        if p and _lib._lib is not None:
            if not _defer("buffer", p, self.nbytes):
                _lib._lib.qa_free(C.c_void_p(p))
            self.ptr = None

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        if arr.nbytes > self.nbytes:
            raise ValueError("upload larger than buffer")
        _lib.call("qa_memcpy_h2d", C.c_void_p(self.ptr), arr.ctypes.data_as(C.c_void_p),
                  C.c_size_t(arr.nbytes), None)

    def download(self, arr):
        if arr.nbytes > self.nbytes:
            raise ValueError("download larger than buffer")
        _lib.call("qa_memcpy_d2h", arr.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr),
                  C.c_size_t(arr.nbytes), None)
        return arr

    @classmethod
    def from_array(cls, arr):
        arr = np.ascontiguousarray(arr)
        b = cls(arr.nbytes)
        b.upload(arr)
        return b


def _vp(x):
    return C.c_void_p(x)


class HipDense:
    """Dense complex128 matrix in HBM, column-major."""

    def __init__(self, shape, buf=None, offset=0):
        self.shape = (int(shape[0]), int(shape[1]))
        n = self.shape[0] * self.shape[1]
        self._buf = buf if buf is not None else DeviceBuffer(16 * n)
        self._offset = offset  # bytes
        self.fortran = True

    # -- qutip Data protocol ---------------------------------------------------
    @classmethod
    def sparcity(cls):
        return "dense"

    @property
    def ptr(self):
        return self._buf.ptr + self._offset

    @property
    def size(self):
        return self.shape[0] * self.shape[1]

    @classmethod
    def from_array(cls, arr, shape=None):
        a = np.asarray(arr, dtype=np.complex128)
        if a.ndim == 1:
            a = a.reshape(-1, 1)
        shape = a.shape if shape is None else shape
        flat = np.ascontiguousarray(a.ravel(order="F"))
        out = cls(shape)
        out._buf.upload(flat)
        return out

    def to_array(self):
        flat = np.empty(self.size, dtype=np.complex128)
        if self.size:
            _lib.call("qa_memcpy_d2h", flat.ctypes.data_as(C.c_void_p), _vp(self.ptr),
                      C.c_size_t(flat.nbytes), None)
        return flat.reshape(self.shape, order="F")

    def copy(self):
        out = HipDense(self.shape)
        _lib.call("qa_memcpy_d2d", _vp(out.ptr), _vp(self.ptr), C.c_size_t(16 * self.size), None)
        return out

    def reshape(self, shape):
        """Same buffer, new column-major shape (stack/unstack columns)."""
        if shape[0] * shape[1] != self.size:
            raise ValueError("reshape changes the number of elements")
        return HipDense(shape, buf=self._buf, offset=self._offset)

    def trace(self):
        if self.shape[0] != self.shape[1]:
            raise ValueError("trace of a non-square matrix")
        res = (C.c_double * 2)()
        _lib.call("qa_trace_z", self.shape[0], _vp(self.ptr), res, None)
        return complex(res[0], res[1])

    def adjoint(self):
        return HipDense.from_array(self.to_array().conj().T)

    def conj(self):
        return HipDense.from_array(self.to_array().conj())

    def transpose(self):
        return HipDense.from_array(self.to_array().T)

    def __reduce__(self):
        return (HipDense.from_array, (self.to_array(),))

    def __repr__(self):
        return f"HipDense(shape={self.shape})"

    # operators route to the dispatchers like qutip's Data (base.pyx:49-93)
    def __add__(self, other):
        return add(self, other)

    def __sub__(self, other):
        return add(self, other, -1)

    def __mul__(self, other):
        return mul(self, other)

    __rmul__ = __mul__

    def __truediv__(self, other):
        return mul(self, 1 / other)

    def __neg__(self):
        return mul(self, -1)

    def __matmul__(self, other):
        return matmul(self, other)


class HipCSR:
    """CSR matrix in HBM (int32 indices, complex128 or complex64 values)."""

    def __init__(self, A, fp32=False):
        A = sp.csr_matrix(A, dtype=np.complex128)
        A.sort_indices()
        self.shape = A.shape
        self.nnz = int(A.nnz)
        self.fp32 = bool(fp32)
        self._host = A
        self._rowptr = DeviceBuffer.from_array(A.indptr.astype(np.int32))
        self._col = DeviceBuffer.from_array(A.indices.astype(np.int32) if self.nnz else
                                            np.zeros(1, np.int32))
        vals = A.data.astype(np.complex64 if fp32 else np.complex128)
        self._vals = DeviceBuffer.from_array(vals if self.nnz else np.zeros(1, vals.dtype))
        self._sell = None

    @classmethod
    def sparcity(cls):
        return "sparse"

    @classmethod
    def from_scipy(cls, A, fp32=False):
        return cls(A, fp32=fp32)

    def to_scipy(self):
        return self._host.copy()

    def to_array(self):
        return self._host.toarray()

    def copy(self):
        return HipCSR(self._host, fp32=self.fp32)

    def adjoint(self):
        return HipCSR(self._host.conj().T, fp32=self.fp32)

    def conj(self):
        return HipCSR(self._host.conj(), fp32=self.fp32)

    def transpose(self):
        return HipCSR(self._host.T, fp32=self.fp32)

    def trace(self):
        return complex(self._host.diagonal().sum())

    def __reduce__(self):
        return (HipCSR, (self._host, self.fp32))

    def __repr__(self):
        return f"HipCSR(shape={self.shape}, nnz={self.nnz}, fp32={self.fp32})"

    def sell(self, lanes=1):
        if self._sell is None or self._sell.lanes != lanes:
            self._sell = SellMatrix(self._host, fp32=self.fp32, lanes=lanes)
        return self._sell

    def __matmul__(self, other):
        return matmul(self, other)


def sell_arrays(A, lanes=1):
    """SELL-C layout of a CSR matrix with ``lanes`` (1, 2, 4, 8) lanes per row
    and C = 64 / lanes rows per slice: per slice the width is the longest
    row rounded up to a multiple of ``lanes``; entry j of row C s + r goes
    to slice_ptr[s] + 64 (j // lanes) + lanes r + j % lanes, so a wavefront
    reads 64 consecutive entries per load.  Padding entries have value 0 and
    the column of their own row (clipped to the column count)."""
    if lanes not in (1, 2, 4, 8):
        raise ValueError("lanes must be 1, 2, 4 or 8")
    A = sp.csr_matrix(A)
    A.sort_indices()
    n = A.shape[0]
    C = 64 // lanes
    ns = (n + C - 1) // C
    npad = ns * C
    rn = np.zeros(npad, dtype=np.int64)
    rn[:n] = np.diff(A.indptr)
    width = rn.reshape(ns, C).max(axis=1)
    width = (width + lanes - 1) // lanes * lanes
    slice_ptr = np.zeros(ns + 1, dtype=np.int64)
    slice_ptr[1:] = np.cumsum(width * C)
    total = int(slice_ptr[-1])
    if total >= 2 ** 31:
        raise ValueError("SELL storage exceeds int32 offsets")
    cols = np.zeros(max(total, 1), dtype=np.int32)
    vals = np.zeros(max(total, 1), dtype=np.complex128)
    rows = np.repeat(np.arange(n), rn[:n])
    j = np.arange(A.nnz) - np.repeat(A.indptr[:-1], rn[:n])
    dest = slice_ptr[rows // C] + (j // lanes) * 64 + (rows % C) * lanes + j % lanes
    cols[dest] = A.indices
    vals[dest] = A.data
    # padding columns: keep loads in range and near the row
    pad_mask = np.ones(total, dtype=bool)
    pad_mask[dest] = False
    if pad_mask.any():
        pidx = np.nonzero(pad_mask)[0]
        s_of = np.searchsorted(slice_ptr, pidx, side="right") - 1
        lane = (pidx - slice_ptr[s_of]) % 64
        r = s_of * C + lane // lanes
        cols[pidx] = np.minimum(r, A.shape[1] - 1).astype(np.int32)
    return slice_ptr.astype(np.int32), cols, vals, npad


CODE_VBITS = range(6, 11)  # value-index bits of a 16-bit code (tables of <= 1024)


def code_split(nvals, noffs):
    """Smallest value-index width b in 6..10 with nvals <= 2^b and
    noffs <= 2^(16 - b), or None."""
    for b in CODE_VBITS:
        if nvals <= (1 << b) and noffs <= (1 << (16 - b)):
            return b
    return None


UNIFORM_WORDS_MAX = 6     # words per lane of the uniform-width coded form
UNIFORM_PAD_MAX = 1.25    # uniform when it stores at most this many times the words


def code_dictionary(A, fp32=False, elem=None):
    """The value and column-offset dictionaries of a sorted CSR matrix:
    (value index per entry, offset index per entry, vtab complex128[2^vbits],
    otab int32[2^(16 - vbits)], vbits, pad code) or None when no 16-bit
    split fits.  Values are told apart by their bits (-0.0 and 0.0 stay
    distinct; ``fp32``: rounded to complex64 first); value 0 and offset 0
    are always present (the padding entry, code ``pad``).

    ``elem`` (one small int per stored entry): the value table is keyed on
    (element, value) and ordered by element, so the entries of element e
    own one contiguous range of the table (element 0 first, with the
    padding value); a seventh item, the range starts per element plus the
    end, is returned."""
    n = A.shape[0]
    vals = A.data.astype(np.complex128)
    if fp32:
        vals = vals.astype(np.complex64).astype(np.complex128)
    bits = np.concatenate([np.zeros((1, 2), np.uint64), vals.view(np.uint64).reshape(-1, 2)])
    if elem is not None:
        el = np.concatenate([np.zeros(1, np.uint64), np.asarray(elem, dtype=np.uint64)])
        bits = np.concatenate([el[:, None], bits], axis=1)
    ubits, vinv = np.unique(bits, axis=0, return_inverse=True)
    rn = np.diff(A.indptr).astype(np.int64)
    rows = np.repeat(np.arange(n, dtype=np.int64), rn)
    offs = np.concatenate([[0], A.indices.astype(np.int64) - rows])
    uoffs, oinv = np.unique(offs, return_inverse=True)
    vbits = code_split(len(ubits), len(uoffs))
    if vbits is None:
        return None
    vinv = vinv.reshape(-1).astype(np.int64)
    oinv = oinv.reshape(-1).astype(np.int64)
    vtab = np.zeros(1 << vbits, dtype=np.complex128)
    vbits_only = np.ascontiguousarray(ubits[:, -2:])
    vtab[:len(ubits)] = vbits_only.view(np.complex128).reshape(-1)
    otab = np.zeros(1 << (16 - vbits), dtype=np.int32)
    otab[:len(uoffs)] = uoffs
    pad = int(vinv[0]) | (int(oinv[0]) << vbits)
    if elem is None:
        return vinv[1:], oinv[1:], vtab, otab, vbits, pad
    nel = int(np.max(elem)) + 1 if len(elem) else 1
    vstart = np.searchsorted(ubits[:, 0], np.arange(nel + 1, dtype=np.uint64)).astype(np.int32)
    return vinv[1:], oinv[1:], vtab, otab, vbits, pad, vstart


def sell_coded_arrays(A, lanes=1, fp32=False, uniform="auto", elem=None):
    """Coded SELL-C layout of a CSR matrix, or None when its distinct values
    and distinct column offsets (column - row) fit no 16-bit split
    (``code_split``: up to 1024 values with 64 offsets ... 64 values with
    1024 offsets).

    The slices and the lane -> entry assignment are those of ``sell_arrays``
    with per-slice widths rounded up to a multiple of 4 * lanes.  Entry j of
    row C s + r is the 16-bit code ``v | (o << vbits)``: v indexes ``vtab``
    (its value, bit-exact; rounded to complex64 first when ``fp32``), o
    indexes ``otab`` (its column minus its row).  A lane's steps 4w .. 4w + 3
    are the four codes of word ``slice_ptr[s] + 64 w + lane`` (little-endian,
    step 4w in the low half of the low 32 bits), so ``slice_ptr`` counts
    8-byte words.  Padding entries are value 0 at offset 0.

    ``uniform``: give every slice the widest slice's width (then slice s
    starts at word s * 64 * words-per-lane, which the stage kernel uses to
    request its codes without reading slice_ptr); "auto" does so when a
    lane holds at most 6 words and the padding stores at most 1.25 times
    the words.

    ``elem``: element of every stored entry (A in the order given, rows
    sorted, duplicates kept), for the time-dependent form: see
    code_dictionary; then the value-table ranges are returned last.

    Returns (slice_ptr int32, codes uint16, vtab complex128[2^vbits],
    otab int32[2^(16 - vbits)], vbits, npad[, vstart])."""
    if lanes not in (1, 2, 4, 8):
        raise ValueError("lanes must be 1, 2, 4 or 8")
    if elem is None:
        A = sp.csr_matrix(A)
        A.sort_indices()
    n = A.shape[0]
    dic = code_dictionary(A, fp32, elem)
    if dic is None:
        return None
    vinv, oinv, vtab, otab, vbits, pad_code = dic[:6]
    rn = np.diff(A.indptr).astype(np.int64)
    rows = np.repeat(np.arange(n, dtype=np.int64), rn)
    C = 64 // lanes
    ns = (n + C - 1) // C
    npad = ns * C
    rnp = np.zeros(npad, dtype=np.int64)
    rnp[:n] = rn
    width = rnp.reshape(ns, C).max(axis=1)
    quantum = 4 * lanes
    width = (width + quantum - 1) // quantum * quantum
    if uniform == "auto":
        wmax = int(width.max()) if len(width) else 0
        uniform = (wmax // lanes // 4 <= UNIFORM_WORDS_MAX
                   and wmax * len(width) <= UNIFORM_PAD_MAX * int(width.sum()))
    if uniform and len(width):
        width = np.full_like(width, width.max())
    words = width * C // 4
    slice_ptr = np.zeros(ns + 1, dtype=np.int64)
    slice_ptr[1:] = np.cumsum(words)
    total_words = int(slice_ptr[-1])
    if total_words >= 2 ** 31:
        raise ValueError("coded SELL storage exceeds int32 word offsets")
    codes = np.full(4 * max(total_words, 1), pad_code, dtype=np.uint16)
    nnz = len(A.data)
    j = np.arange(nnz, dtype=np.int64) - np.repeat(A.indptr[:-1].astype(np.int64), rn)
    u, q = j // lanes, j % lanes
    lane = (rows % C) * lanes + q
    word = slice_ptr[rows // C] + (u // 4) * 64 + lane
    codes[4 * word + u % 4] = (vinv | (oinv << vbits)).astype(np.uint16)
    out = (slice_ptr.astype(np.int32), codes, vtab, otab, vbits, npad)
    return out if elem is None else out + (dic[6],)


def stacked_elements(mats):
    """The stored entries of several n x n CSR matrices (the elements of a
    time-dependent generator, element 0 the constant part) as one CSR whose
    rows hold every element's entries, duplicates kept (element by element,
    columns sorted inside each), and the element of every entry."""
    mats = [sp.csr_matrix(M) for M in mats]
    n = mats[0].shape[0]
    rows, cols, vals, elem = [], [], [], []
    for e, M in enumerate(mats):
        M = M.copy()
        M.sum_duplicates()
        M.sort_indices()
        rn = np.diff(M.indptr)
        rows.append(np.repeat(np.arange(n, dtype=np.int64), rn))
        cols.append(M.indices.astype(np.int64))
        vals.append(M.data.astype(np.complex128))
        elem.append(np.full(M.nnz, e, dtype=np.int64))
    rows, cols, vals, elem = map(np.concatenate, (rows, cols, vals, elem))
    order = np.lexsort((cols, elem, rows))
    rows, cols, vals, elem = rows[order], cols[order], vals[order], elem[order]
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(indptr, rows + 1, 1)
    indptr = np.cumsum(indptr)
    A = sp.csr_matrix((vals, cols.astype(np.int32), indptr.astype(np.int32)), shape=(n, n))
    return A, elem


class SellMatrix:
    """A SELL-C copy of a CSR matrix in HBM (see sell_arrays), or with
    ``coded=True`` the coded form (sell_coded_arrays; ValueError when the
    matrix does not fit the 256-entry tables).  ``elem``: the element of
    every stored entry of a stacked time-dependent generator
    (stacked_elements); the coded value table is then split by element and
    ``vstart`` holds the ranges."""

    def __init__(self, A, fp32=False, lanes=1, coded=False, uniform="auto", elem=None):
        self.shape = A.shape
        self.lanes = lanes
        self.fp32 = fp32
        self.coded = bool(coded) or elem is not None
        self.vstart = None
        if self.coded:
            arrs = sell_coded_arrays(A, lanes, fp32, uniform, elem=elem)
            if arrs is None:
                raise ValueError("the operator's distinct values and column offsets fit no "
                                 "16-bit code split: no coded SELL-C form")
            if elem is not None:
                self.vstart = arrs[-1]
                arrs = arrs[:-1]
            sptr, codes, vtab, otab, self.vbits, npad = arrs
            self.nslices = len(sptr) - 1
            self.npad = npad
            self.stored = 4 * int(sptr[-1])
            self.uniform = len(set(np.diff(sptr).tolist())) == 1
            self._sptr = DeviceBuffer.from_array(sptr)
            self._cols = DeviceBuffer.from_array(codes)
            self._vals = DeviceBuffer.from_array(vtab)
            self._otab = DeviceBuffer.from_array(otab)
            return
        sptr, cols, vals, npad = sell_arrays(A, lanes)
        self.nslices = len(sptr) - 1
        self.npad = npad
        self.stored = int(sptr[-1])
        self._sptr = DeviceBuffer.from_array(sptr)
        self._cols = DeviceBuffer.from_array(cols)
        self._vals = DeviceBuffer.from_array(vals.astype(np.complex64 if fp32 else np.complex128))

    @property
    def ptrs(self):
        """(slice_ptr, columns, values); coded: (slice_ptr, codes, value table)."""
        return self._sptr.ptr, self._cols.ptr, self._vals.ptr

    @property
    def otab_ptr(self):
        return self._otab.ptr

    def matvec(self, x, out, scale=1.0, accumulate=False):
        sre, sim = _scalar(scale)
        if self.coded:
            _lib.call("qa_sell_coded_matvec_z", self.shape[0], self.shape[1], self.lanes,
                      self.nslices, _vp(self._sptr.ptr), _vp(self._cols.ptr),
                      _vp(self._vals.ptr), _vp(self._otab.ptr), self.vbits, _vp(x.ptr),
                      _vp(out.ptr), sre, sim, int(accumulate), None)
            return out
        _lib.call("qa_sell_matvec_z", self.shape[0], self.lanes, self.nslices,
                  _vp(self._sptr.ptr), _vp(self._cols.ptr), _vp(self._vals.ptr), int(self.fp32),
                  _vp(x.ptr), _vp(out.ptr), sre, sim, int(accumulate), None)
        return out


# ---------------------------------------------------------------------------
# conversions (the `to` graph edges Dense <-> HipDense, CSR <-> HipCSR)
# ---------------------------------------------------------------------------
def to_hip(x, fp32=False):
    if isinstance(x, (HipDense, HipCSR)):
        return x
    if sp.issparse(x):
        return HipCSR(x, fp32=fp32)
    return HipDense.from_array(np.asarray(x))


def to_numpy(x):
    if isinstance(x, (HipDense, HipCSR)):
        return x.to_array()
    return np.asarray(x)


# ---------------------------------------------------------------------------
# dispatcher specialisations
# ---------------------------------------------------------------------------
def _scalar(s):
    s = complex(s)
    return s.real, s.imag


def zeros(rows, cols):
    """zeros specialisation (qutip/core/data/constant.py:18-45)."""
    out = HipDense((rows, cols))
    _lib.call("qa_fill_zero", _vp(out.ptr), C.c_size_t(16 * out.size), None)
    return out


def matmul(left, right, scale=1, out=None):
    """matmul(left, right, scale=1) -> HipDense (qutip/core/data/matmul.pyx:1018-1029);
    with ``out`` it is the in-place ``out += scale * left @ right`` of
    imatmul_data_dense (matmul.pyx:1207-1215)."""
    if not isinstance(right, HipDense):
        raise TypeError(f"matmul: right operand must be HipDense, got {type(right).__name__}")
    if left.shape[1] != right.shape[0]:
        raise ValueError(f"incompatible matrix shapes {left.shape} and {right.shape}")
    acc = out is not None
    if out is None:
        out = HipDense((left.shape[0], right.shape[1]))
    elif out.shape != (left.shape[0], right.shape[1]):
        raise ValueError("out has the wrong shape")
    sre, sim = _scalar(scale)
    if isinstance(left, HipCSR):
        if right.shape[1] == 1:
            _lib.call("qa_csr_matvec_z", left.shape[0], left.shape[1], _vp(left._rowptr.ptr),
                      _vp(left._col.ptr), _vp(left._vals.ptr), int(left.fp32), left.nnz,
                      _vp(right.ptr), _vp(out.ptr), sre, sim, int(acc), None)
        else:
            # several columns: the SELL-C copy (built once per operator), each
            # stored entry read once per 8 columns
            S = left.sell(1)
            sptr, cols, vals = S.ptrs
            _lib.call("qa_sell_matmat_z", left.shape[0], left.shape[1], S.lanes, S.nslices,
                      _vp(sptr), _vp(cols), _vp(vals), int(left.fp32), right.shape[1],
                      _vp(right.ptr), right.shape[0], _vp(out.ptr), out.shape[0], sre, sim,
                      int(acc), None)
        return out
    raise TypeError(f"matmul: no device specialisation for {type(left).__name__} @ HipDense")


def add(left, right, scale=1):
    """left + scale * right (qutip/core/data/add.pyx:324-374)."""
    if left.shape != right.shape:
        raise ValueError(f"incompatible matrix shapes {left.shape} and {right.shape}")
    out = HipDense(left.shape)
    sre, sim = _scalar(scale)
    _lib.call("qa_axpby_z", left.size, 1.0, 0.0, _vp(left.ptr), sre, sim, _vp(right.ptr),
              _vp(out.ptr), None)
    return out


def iadd(left, right, scale=1):
    """In place: left += scale * right (returns left)."""
    if left.shape != right.shape:
        raise ValueError(f"incompatible matrix shapes {left.shape} and {right.shape}")
    sre, sim = _scalar(scale)
    _lib.call("qa_axpby_z", left.size, 1.0, 0.0, _vp(left.ptr), sre, sim, _vp(right.ptr),
              _vp(left.ptr), None)
    return left


def mul(matrix, value):
    """value * matrix (qutip/core/data/mul.pyx:103-139)."""
    if not isinstance(value, numbers.Number):
        raise TypeError("mul: value must be a scalar")
    out = HipDense(matrix.shape)
    vre, vim = _scalar(value)
    _lib.call("qa_axpby_z", matrix.size, vre, vim, _vp(matrix.ptr), 0.0, 0.0, None,
              _vp(out.ptr), None)
    return out


def imul(matrix, value):
    vre, vim = _scalar(value)
    _lib.call("qa_axpby_z", matrix.size, vre, vim, _vp(matrix.ptr), 0.0, 0.0, None,
              _vp(matrix.ptr), None)
    return matrix


def norm_l2(x):
    """||x||_2 of a ket / bra (qutip/core/data/norm.pyx:127-135)."""
    if x.shape[1] != 1 and x.shape[0] != 1:
        raise ValueError("l2 norm is only defined for vectors")
    return frobenius(x)


def frobenius(x):
    r = C.c_double()
    _lib.call("qa_norm2_z", x.size, _vp(x.ptr), C.byref(r), None)
    return float(np.sqrt(r.value))


def wrmn_error(diff, state, atol, rtol):
    """Weighted RMS error norm (qutip/core/data/ode.pyx:167-207)."""
    if diff.shape != state.shape:
        raise ValueError(f"incompatible matrix shapes {diff.shape} and {state.shape}")
    r = C.c_double()
    _lib.call("qa_wrmn_error_z", diff.size, _vp(diff.ptr), _vp(state.ptr), float(atol),
              float(rtol), C.byref(r), None)
    return r.value


def inner(left, right, scalar_is_ket=False):
    """<left|right> for kets (qutip/core/data/inner.pyx:323-334): left is
    treated as a ket (conjugated) when its shape is a column."""
    if left.size != right.size:
        raise ValueError(f"incompatible matrix shapes {left.shape} and {right.shape}")
    res = (C.c_double * 2)()
    if left.shape[0] == 1 and left.shape[1] != 1:
        # a bra: <bra|ket> = sum bra_i ket_i  (no conjugation)
        conj = HipDense.from_array(left.to_array().conj())
        _lib.call("qa_inner_z", left.size, _vp(conj.ptr), _vp(right.ptr), res, None)
    else:
        _lib.call("qa_inner_z", left.size, _vp(left.ptr), _vp(right.ptr), res, None)
    return complex(res[0], res[1])


def expect(op, state):
    """expect(op, state) -> complex (qutip/core/data/expect.pyx:391-400):
    <psi|op|psi> for a ket, tr(op rho) for a square state."""
    if not isinstance(op, HipCSR):
        op = HipCSR(sp.csr_matrix(to_numpy(op)))
    res = (C.c_double * 2)()
    if state.shape[1] == 1:
        if op.shape[1] != state.shape[0]:
            raise ValueError("incompatible shapes for expect")
        _lib.call("qa_expect_csr_ket_z", op.shape[0], _vp(op._rowptr.ptr), _vp(op._col.ptr),
                  _vp(op._vals.ptr), int(op.fp32), _vp(state.ptr), res, None)
    else:
        if op.shape[1] != state.shape[0] or state.shape[0] != state.shape[1]:
            raise ValueError("incompatible shapes for expect")
        _lib.call("qa_expect_csr_dm_z", op.shape[0], _vp(op._rowptr.ptr), _vp(op._col.ptr),
                  _vp(op._vals.ptr), int(op.fp32), _vp(state.ptr), res, None)
    return complex(res[0], res[1])


def trace(x):
    return x.trace()


def column_stack(x):
    """vec(rho): free for column-major storage (reshape.pyx:208-286)."""
    return x.reshape((x.size, 1))


def column_unstack(x, rows):
    if x.size % rows:
        raise ValueError("rows does not divide the vector length")
    return x.reshape((rows, x.size // rows))
