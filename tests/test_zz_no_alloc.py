"""The allocation-free error norm qa_wrmn_error_z_noalloc through the C ABI:
the same bits as qa_wrmn_error_z (same kernels and grid), on two calls in a
row, the second with a smaller input, and the closed form
sqrt(mean((|d| / (atol + rtol |s|))^2)) of wrmn_error_dense
(qutip/core/data/ode.pyx:27-52)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from qutip_amd import _lib  # noqa: E402


def _device(arr):
    arr = np.ascontiguousarray(arr, dtype=np.complex128)
    p = C.c_void_p()
    _lib.call("qa_malloc", C.byref(p), arr.nbytes)
    _lib.call("qa_memcpy_h2d", p, arr.ctypes.data_as(C.c_void_p), arr.nbytes, None)
    return p


def _norm(name, n, d, s, atol, rtol):
    out = C.c_double(-1.0)
    _lib.call(name, n, d, s, atol, rtol, C.byref(out), None)
    return out.value


@pytest.mark.parametrize("sizes", [(1_000_003, 4097), (300_000, 1)])
def test_noalloc_equals_wrmn_error(sizes):
    rng = np.random.default_rng(5)
    atol, rtol = 1e-8, 1e-6
    for n in sizes:
        print(f"wrmn_error noalloc n={n}", flush=True)
        d = rng.standard_normal(n) * 1e-7 + 1j * rng.standard_normal(n) * 1e-7
        s = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        D, S = _device(d), _device(s)
        try:
            want = _norm("qa_wrmn_error_z", n, D, S, atol, rtol)
            got = _norm("qa_wrmn_error_z_noalloc", n, D, S, atol, rtol)
            again = _norm("qa_wrmn_error_z_noalloc", n, D, S, atol, rtol)
        finally:
            _lib.call("qa_free", D)
            _lib.call("qa_free", S)
        closed = np.sqrt(np.mean((np.abs(d) / (atol + rtol * np.abs(s))) ** 2))
        assert got == want and again == want, (got, again, want)
        assert got == pytest.approx(closed, rel=1e-12)


def test_noalloc_rejects_empty_input():
    with pytest.raises(ValueError):
        _norm("qa_wrmn_error_z_noalloc", 0, None, None, 1e-8, 1e-6)
