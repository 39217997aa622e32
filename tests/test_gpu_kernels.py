"""Data-layer kernels on the GPU against numpy (tolerances of
qutip/tests/core/data/test_mathematics.py:369-370: atol 1e-10, rtol 1e-7),
mirroring TestMatmul / TestInner / TestExpect of the reference suite."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
from numpy.testing import assert_allclose

pytestmark = pytest.mark.gpu

from qutip_amd import _lib, data as qd  # noqa: E402

ATOL, RTOL = 1e-10, 1e-7


def rand_csr(n, m, density, seed, shuffle=False):
    rng = np.random.default_rng(seed)
    A = sp.random(n, m, density=density, random_state=rng, format="csr")
    A = sp.csr_matrix(A + 1j * sp.random(n, m, density=density, random_state=rng, format="csr"))
    return A


def rand_vec(n, seed, cols=1):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, cols)) + 1j * rng.standard_normal((n, cols))


@pytest.mark.parametrize("shape,density", [((100, 100), 0.1), ((100, 37), 0.3), ((1, 100), 0.5),
                                           ((100, 1), 0.5), ((257, 257), 0.02), ((50, 50), 0.0),
                                           ((2000, 2000), 0.01)])
@pytest.mark.parametrize("fp32", [False, True])
def test_csr_matvec(shape, density, fp32):
    print(f"csr_matvec shape={shape} density={density} fp32={fp32}", flush=True)
    A = rand_csr(*shape, density, seed=shape[0])
    x = rand_vec(shape[1], 1)
    Ad = qd.HipCSR(A, fp32=fp32)
    got = qd.matmul(Ad, qd.HipDense.from_array(x), scale=0.5 - 2j).to_array()
    Aref = A.astype(np.complex64).astype(np.complex128) if fp32 else A
    assert_allclose(got, (0.5 - 2j) * (Aref @ x), atol=ATOL, rtol=RTOL)


def test_csr_matvec_inplace_accumulates():
    A = rand_csr(300, 300, 0.05, 3)
    x, y = rand_vec(300, 4), rand_vec(300, 5)
    out = qd.HipDense.from_array(y)
    qd.matmul(qd.HipCSR(A), qd.HipDense.from_array(x), scale=1j, out=out)
    assert_allclose(out.to_array(), y + 1j * (A @ x), atol=ATOL, rtol=RTOL)


def test_csr_matmat_columns():
    A = rand_csr(120, 90, 0.1, 6)
    X = rand_vec(90, 7, cols=5)
    got = qd.matmul(qd.HipCSR(A), qd.HipDense.from_array(X)).to_array()
    assert_allclose(got, A @ X, atol=ATOL, rtol=RTOL)


@pytest.mark.parametrize("lanes", [1, 2, 4, 8])
@pytest.mark.parametrize("n", [1, 64, 65, 1000, 65536])
def test_sell_matvec(n, lanes):
    print(f"sell_matvec n={n} lanes={lanes}", flush=True)
    A = rand_csr(n, n, min(1.0, 12.0 / n), n)
    x = rand_vec(n, 2)
    S = qd.SellMatrix(A, lanes=lanes)
    out = S.matvec(qd.HipDense.from_array(x), qd.HipDense((n, 1)))
    assert_allclose(out.to_array(), A @ x, atol=ATOL, rtol=RTOL)


def test_mismatched_shapes_raise():
    A = qd.HipCSR(rand_csr(10, 10, 0.3, 1))
    with pytest.raises(ValueError):
        qd.matmul(A, qd.HipDense.from_array(rand_vec(9, 1)))


def test_add_mul_norm_inner_wrmn():
    x, y = rand_vec(1000, 1), rand_vec(1000, 2)
    X, Y = qd.HipDense.from_array(x), qd.HipDense.from_array(y)
    assert_allclose(qd.add(X, Y, 2 - 1j).to_array(), x + (2 - 1j) * y, atol=ATOL, rtol=RTOL)
    assert_allclose(qd.mul(X, 0.25j).to_array(), 0.25j * x, atol=ATOL, rtol=RTOL)
    assert qd.frobenius(X) == pytest.approx(np.linalg.norm(x), rel=1e-12)
    assert qd.inner(X, Y) == pytest.approx(np.vdot(x, y), rel=1e-12)
    want = np.sqrt(np.mean((np.abs(x) / (1e-3 + 1e-2 * np.abs(y))) ** 2))
    assert qd.wrmn_error(X, Y, 1e-3, 1e-2) == pytest.approx(want, rel=1e-12)
    qd.iadd(X, Y, -1)
    assert_allclose(X.to_array(), x - y, atol=ATOL, rtol=RTOL)


def test_expect_ket_and_dm():
    d = 40
    O = rand_csr(d, d, 0.2, 9)
    psi = rand_vec(d, 3)
    got = qd.expect(qd.HipCSR(O), qd.HipDense.from_array(psi))
    assert got == pytest.approx(complex(np.vdot(psi, O @ psi)), rel=1e-10)
    r = rand_vec(d, 4, cols=d)
    rho = r @ r.conj().T
    got = qd.expect(qd.HipCSR(O), qd.HipDense.from_array(rho))
    assert got == pytest.approx(complex(np.trace(O @ rho)), rel=1e-10)
    assert qd.HipDense.from_array(rho).trace() == pytest.approx(np.trace(rho), rel=1e-12)


def test_zeros_and_copy_roundtrip():
    z = qd.zeros(17, 3)
    assert not z.to_array().any()
    x = rand_vec(51, 8, cols=2)
    X = qd.HipDense.from_array(x)
    assert np.array_equal(X.copy().to_array(), x)


@pytest.mark.parametrize("shape,density", [((1000, 1000), 0.01), ((300, 77), 0.2),
                                           ((65, 65), 0.5), ((40, 40), 0.0)])
@pytest.mark.parametrize("ncols", [2, 8, 13])
@pytest.mark.parametrize("fp32", [False, True])
def test_sell_matmat_against_scipy(shape, density, ncols, fp32):
    """data.matmul with several right-hand columns runs the SELL-C block
    product (each stored entry read once per 8 columns): against scipy,
    with a complex scale and accumulation into out."""
    print(f"sell_matmat shape={shape} density={density} ncols={ncols} fp32={fp32}", flush=True)
    A = rand_csr(*shape, density, seed=shape[0] + ncols)
    X = rand_vec(shape[1], 3, ncols)
    Ad = qd.HipCSR(A, fp32=fp32)
    Aref = A.astype(np.complex64).astype(np.complex128) if fp32 else A
    got = qd.matmul(Ad, qd.HipDense.from_array(X), scale=0.5 - 2j).to_array()
    assert_allclose(got, (0.5 - 2j) * (Aref @ X), atol=ATOL, rtol=RTOL)
    base = rand_vec(shape[0], 4, ncols)
    out = qd.HipDense.from_array(base)
    qd.matmul(Ad, qd.HipDense.from_array(X), scale=1.5, out=out)
    assert_allclose(out.to_array(), base + 1.5 * (Aref @ X), atol=ATOL, rtol=RTOL)


@pytest.mark.parametrize("nvec", [1, 8, 13])
def test_csr_matmat_abi_column_tiles(nvec):
    """qa_csr_matmat_z (one thread per row and 8 columns) with leading
    dimensions larger than the matrix."""
    n, m = 513, 300
    A = rand_csr(n, m, 0.05, seed=nvec)
    X = np.zeros((m + 5, nvec), dtype=np.complex128)
    X[:m] = rand_vec(m, 5, nvec)
    Ad = qd.HipCSR(A)
    xd = qd.HipDense.from_array(np.asfortranarray(X))
    out = qd.HipDense((n + 3, nvec))
    _lib.call("qa_csr_matmat_z", n, m, C.c_void_p(Ad._rowptr.ptr), C.c_void_p(Ad._col.ptr),
              C.c_void_p(Ad._vals.ptr), 0, nvec, C.c_void_p(xd.ptr), m + 5, C.c_void_p(out.ptr),
              n + 3, 1.0, 0.0, 0, None)
    assert_allclose(out.to_array()[:n], A @ X[:m], atol=ATOL, rtol=RTOL)
