"""The matrix-form Lindblad path (MFMA products and the FP64 banded form) on
the GPU.

* LindbladMatrixForm.matmul_data against the dense formula in numpy (FP32
  and bf16 x6 products: relative error ~1e-6 of the operator norms; bf16 x3
  keeps ~16 mantissa bits per factor);
* mesolve(matrix_form=True) against the superoperator path and the FP64
  fixture of the driven Kerr oscillator (N=256) over the whole window
  t in [0, 30]: FP32, bf16 x6 and the banded form within 1e-6, bf16 x3
  within its stated 2e-4;
* the banded form (FP64 products of the operators' diagonals) against the
  numpy formula to rounding (1e-13 relative, ragged N) and the Kerr fixture
  to 1e-6."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import qutip_amd as qa  # noqa: E402
from tests import systems  # noqa: E402


def _dense_rhs(H, cs, rho):
    Hnh = H - 0.5j * sum(c.conj().T @ c for c in cs)
    out = -1j * Hnh @ rho + 1j * rho @ Hnh.conj().T
    for c in cs:
        out += c @ rho @ c.conj().T
    return out


@pytest.mark.parametrize("N", [16, 32, 64, 128, 384, 512])
@pytest.mark.parametrize("prec", ["fp32", "bf16x6", "bf16x3"])
def test_matrix_form_rhs_vs_numpy(N, prec):
    # N >= 384 runs the 32 x 32 LDS-tiled kernels, smaller N the one-tile kernels
    if prec.startswith("bf16") and N % 32:
        pytest.skip("the bf16 forms need N % 32 == 0")
    print(f"dense rhs N={N} prec={prec}", flush=True)
    rng = np.random.default_rng(N)
    H = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    H = (H + H.conj().T) / 2
    cs = [(rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))) * 0.3 for _ in range(2)]
    r = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    rho = r @ r.conj().T / N
    L = qa.LindbladMatrixForm(qa.Qobj(H), [qa.Qobj(c) for c in cs], precision=prec)
    got = L.matmul_data(0.0, qa.HipDense.from_array(rho)).to_array()
    want = _dense_rhs(H, cs, rho)
    rel = np.abs(got - want).max() / np.abs(want).max()
    print(f"  relative max error {rel:.2e}", flush=True)
    # FP32 and bf16 x6: 24-bit inputs, FP32 sums; bf16 x3 keeps ~16 bits
    # of each input (its dropped lo*lo term is 2^-16 relative)
    assert rel < {"fp32": 2e-5, "bf16x6": 2e-5, "bf16x3": 1e-4}[prec]


def test_bf16x6_is_as_accurate_as_fp32():
    """The three-piece split reproduces the FP32 inputs exactly, so its RHS
    error against complex128 is that of FP32 (within a factor for the
    different FP32 summation order), and 100x below bf16 x3's."""
    N = 256
    rng = np.random.default_rng(3)
    H = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    H = (H + H.conj().T) / 2
    cs = [(rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))) * 0.3]
    r = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    rho = r @ r.conj().T / N
    want = _dense_rhs(H, cs, rho)
    rel = {}
    for prec in ("fp32", "bf16x6", "bf16x3"):
        L = qa.LindbladMatrixForm(qa.Qobj(H), [qa.Qobj(c) for c in cs], precision=prec)
        got = L.matmul_data(0.0, qa.HipDense.from_array(rho)).to_array()
        rel[prec] = np.abs(got - want).max() / np.abs(want).max()
    print(f"relative RHS errors {rel}", flush=True)
    assert rel["bf16x6"] < 3 * rel["fp32"]
    assert rel["bf16x6"] * 30 < rel["bf16x3"]


# the metric's bar: <O>(t) within 1e-6 of the FP64 solution over the whole
# window (t in [0, 30], 150 points); bf16 x3 is outside it and stated
KERR_BAR = {"fp32": 1e-6, "bf16x6": 1e-6, "bf16x3": 2e-4, "fp64_banded": 1e-6}


@pytest.mark.parametrize("prec", ["fp32", "bf16x6", "bf16x3", "fp64_banded"])
def test_kerr_matrix_form_full_window(golden, prec):
    print(f"kerr matrix form prec={prec}, t in [0, 30]", flush=True)
    g = golden("kerr_full_mesolve.npz")
    s = systems.kerr(256)
    res = qa.mesolve(s["H"], s["psi0"], g["tlist"], s["c_ops"], e_ops=s["e_ops"],
                     options={"matrix_form": True, "dense_precision": prec,
                              "rtol": 1e-8, "atol": 1e-10})
    errs = [np.abs(np.asarray(res.expect[k]) - g["expect"][k]).max() for k in range(2)]
    print(f"  <n> err {errs[0]:.2e} (max <n> {g['expect'][0].max():.2f}), parity err "
          f"{errs[1]:.2e}; stats {res.stats}", flush=True)
    assert max(errs) < KERR_BAR[prec]


def test_matrix_form_matches_superoperator_small():
    """A 16-level damped driven oscillator: matrix form (FP32 MFMA) against
    the FP64 superoperator path."""
    N = 16
    a = qa.destroy(N)
    H = 0.5 * a.dag() @ a + 0.3 * (a + a.dag())
    c_ops = [np.sqrt(0.4) * a]
    tlist = np.linspace(0, 5, 26)
    e = [a.dag() @ a]
    ref = qa.mesolve(H, qa.basis(N, 0), tlist, c_ops, e_ops=e, options={"rtol": 1e-8, "atol": 1e-10})
    got = qa.mesolve(H, qa.basis(N, 0), tlist, c_ops, e_ops=e,
                     options={"matrix_form": True, "rtol": 1e-8, "atol": 1e-10})
    err = np.abs(np.asarray(got.expect[0]) - np.asarray(ref.expect[0])).max()
    print(f"matrix form vs superoperator: {err:.2e}", flush=True)
    assert err < 1e-5


@pytest.mark.parametrize("prec", ["fp32", "bf16x6", "bf16x3"])
def test_matrix_form_decay_closed_form_large(prec):
    """The whole tiled stage path (N = 384: 32 x 32 LDS tiles) on a solve with
    a closed form: H = 0, c = sqrt(g) a, rho0 = |5><5| gives
    <a^dag a>(t) = 5 exp(-g t) (qutip/tests/solver/test_mesolve.py:84-95
    uses the same decay)."""
    N, g = 384, 0.4
    a = qa.destroy(N)
    H = qa.Qobj(np.zeros((N, N), dtype=complex))
    tl = np.linspace(0, 4, 9)
    res = qa.mesolve(H, qa.basis(N, 5), tl, [np.sqrt(g) * a], e_ops=[a.dag() @ a],
                     options={"matrix_form": True, "dense_precision": prec, "rtol": 1e-8,
                              "atol": 1e-10})
    err = np.abs(np.asarray(res.expect[0]) - 5 * np.exp(-g * tl)).max()
    print(f"decay N={N} prec={prec}: max err {err:.2e}", flush=True)
    assert err < (5e-4 if prec == "bf16x3" else 1e-5)


def _banded_system(N, rng):
    """Hermitian H on diagonals 0, +-1, +-3; c1 on diagonal +1 (a lowering
    ladder), c2 on diagonals 0 and -2."""
    H = np.zeros((N, N), complex)
    for o, scale in ((0, 1.0), (1, 0.5), (3, 0.2)):
        v = (rng.standard_normal(N - o) + 1j * rng.standard_normal(N - o)) * scale
        H += np.diag(v, o) + (np.diag(v.conj(), -o) if o else 0)
    H = (H + H.conj().T) / 2
    c1 = np.diag(np.sqrt(np.arange(1, N)) * 0.3, 1).astype(complex)
    c2 = np.diag(rng.standard_normal(N) * 0.2 + 0j) + np.diag(rng.standard_normal(N - 2) * 0.1j, -2)
    return H, [c1, c2]


@pytest.mark.parametrize("N", [37, 64, 256])
def test_banded_rhs_vs_numpy(N):
    """The FP64 banded form (products of diagonals, any N) against the dense
    formula in numpy: rounding-level agreement."""
    print(f"banded rhs N={N}", flush=True)
    rng = np.random.default_rng(N + 1)
    H, cs = _banded_system(N, rng)
    r = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    rho = r @ r.conj().T / N
    L = qa.LindbladMatrixForm(qa.Qobj(H), [qa.Qobj(c) for c in cs], precision="fp64_banded")
    got = L.matmul_data(0.0, qa.HipDense.from_array(rho)).to_array()
    assert L._engine.dense_precision == "fp64_banded"
    want = _dense_rhs(H, cs, rho)
    rel = np.abs(got - want).max() / np.abs(want).max()
    print(f"  relative max error {rel:.2e}", flush=True)
    assert rel < 1e-13


def test_kerr_banded_vs_golden(golden):
    """Kerr N=256 (H_nh on 3 diagonals, a on one): 'auto' takes the banded
    FP64 form; <n> and parity within 1e-6 of the FP64 fixture."""
    print("kerr matrix form banded", flush=True)
    g = golden("kerr_mesolve.npz")
    s = systems.kerr(256)
    solver = qa.MESolver(s["H"], s["c_ops"], options={"matrix_form": True, "rtol": 1e-8,
                                                      "atol": 1e-10})
    res = solver.run(s["psi0"], g["tlist"], e_ops=s["e_ops"])
    assert solver._integrator._engine.dense_precision == "fp64_banded"
    errs = [np.abs(np.asarray(res.expect[k]) - g["expect"][k]).max() for k in range(2)]
    print(f"  <n> err {errs[0]:.2e}, parity err {errs[1]:.2e}; stats {res.stats}", flush=True)
    assert max(errs) < 1e-6


def test_auto_keeps_mfma_for_dense_operators():
    """A dense collapse operator has N diagonals: 'auto' takes the MFMA
    form."""
    N = 32
    rng = np.random.default_rng(3)
    H, _ = _banded_system(N, rng)
    c = (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))) * 0.1
    L = qa.LindbladMatrixForm(qa.Qobj(H), [qa.Qobj(c)])
    r = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    L.matmul_data(0.0, qa.HipDense.from_array(r @ r.conj().T / N))
    assert L._engine.dense_precision == "fp32"
