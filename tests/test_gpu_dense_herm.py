"""The matrix-form engine with Hermitian states (assume_hermitian_state,
qutip/core/cy/lindblad_matrix_form.pyx:59-90 and :175-189): each stage
computes the tiles on and above the diagonal and writes them mirrored.

* against the full-RHS kernels on the same solve: <O>(t) within the FP32
  product's rounding, for N on both kernel paths (16 x 16 tiles below 384,
  LDS-staged 32 x 32 tiles from 384) and every MFMA form;
* the state after the solve is exactly Hermitian (bit for bit), as the
  reference's A + A^dag makes every RHS;
* MESolver switches the fast path off for an operator state that is not
  Hermitian (mesolve.py:270-276) and that solve matches the FP64
  superoperator path."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import qutip_amd as qa  # noqa: E402
from qutip_amd.solver.hip_rk import RKEngine  # noqa: E402


def _system(N, seed, nc=2):
    rng = np.random.default_rng(seed)
    H = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    H = (H + H.conj().T) / (2 * np.sqrt(N))
    cs = [(rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))) * (0.3 / np.sqrt(N))
          for _ in range(nc)]
    Hnh = H - 0.5j * sum(c.conj().T @ c for c in cs)
    r = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    rho0 = r @ r.conj().T
    rho0 /= np.trace(rho0).real
    return H, cs, Hnh, rho0


def _trace_functionals(ops, N):
    """tr(O rho) for each O, rho column-stacked: rho[c, r] at r N + c."""
    kinds, fptr, rows, cols, ws = [], [0], [], [], []
    for O in ops:
        r, c = np.nonzero(O)
        kinds.append(0)
        rows.append(np.zeros(len(r), np.int32))
        cols.append((r * N + c).astype(np.int32))
        ws.append(O[r, c].astype(np.complex128))
        fptr.append(fptr[-1] + len(r))
    return (np.array(kinds, np.int32), np.array(fptr, np.int32), np.concatenate(rows),
            np.concatenate(cols), np.concatenate(ws))


def _solve(Hnh, cs, rho0, tl, prec, herm, ops):
    eng = RKEngine(None, "vern7", 1e-8, 1e-10, dense=(Hnh, cs), dense_precision=prec)
    eng.set_hermitian(herm)
    assert eng.hermitian == herm
    eng.set_functionals(*_trace_functionals(ops, Hnh.shape[0]))
    eng.set_state(tl[0], rho0.ravel(order="F"))
    vals, _, _ = eng.run(tl[1:])
    _, y = eng.get_state()
    return vals, y.to_array().reshape(Hnh.shape, order="F"), eng.stats()


@pytest.mark.parametrize("N", [32, 96, 128, 256, 384, 512])
@pytest.mark.parametrize("prec", ["fp32", "bf16x6", "bf16x3"])
def test_upper_tiles_match_full_rhs(N, prec):
    print(f"hermitian dense N={N} prec={prec}", flush=True)
    H, cs, Hnh, rho0 = _system(N, 10 + N)
    tl = np.linspace(0, 0.5, 6)
    ops = [H, np.diag(np.arange(N) / N).astype(complex)]
    vh, yh, sh = _solve(Hnh, cs, rho0, tl, prec, True, ops)
    vf, yf, sf = _solve(Hnh, cs, rho0, tl, prec, False, ops)
    diff = np.abs(vh - vf).max()
    scale = max(1.0, np.abs(vf).max())
    print(f"  max |herm - full| {diff:.2e} (scale {scale:.2f}); attempts {sh['attempts']} / "
          f"{sf['attempts']}; herm state anti-Hermitian part "
          f"{np.abs(yh - yh.conj().T).max():.1e}", flush=True)
    # the two kernels sum the FP32 products in different orders; bf16 x3's
    # inputs keep 16 bits
    assert diff < (2e-4 if prec == "bf16x3" else 2e-6) * scale
    assert np.array_equal(yh, yh.conj().T), "the Hermitian path's state is not exactly Hermitian"
    assert np.all(yh.diagonal().imag == 0.0)


def test_non_hermitian_state_takes_full_rhs():
    """rho0 = |0><1| + its Hermitian part: not Hermitian, so MESolver turns
    assume_hermitian_state off and the engine runs the full RHS; <O>(t)
    matches the FP64 superoperator path."""
    N = 64
    a = qa.destroy(N)
    H = 0.5 * a.dag() @ a + 0.2 * (a + a.dag())
    c_ops = [np.sqrt(0.3) * a]
    m = np.zeros((N, N), complex)
    m[0, 0] = 0.5
    m[1, 1] = 0.5
    m[0, 1] = 0.3
    rho0 = qa.Qobj(m)
    assert not rho0.isherm
    tl = np.linspace(0, 2, 11)
    e = [a, a.dag() @ a]
    solver = qa.MESolver(H, c_ops, options={"matrix_form": True, "dense_precision": "fp32",
                                            "rtol": 1e-8, "atol": 1e-10})
    got = solver.run(rho0, tl, e_ops=e)
    assert solver.rhs.assume_hermitian_state is False
    assert solver._integrator._engine.hermitian is False
    ref = qa.mesolve(H, rho0, tl, c_ops, e_ops=e, options={"rtol": 1e-8, "atol": 1e-10})
    err = max(np.abs(np.asarray(got.expect[k]) - np.asarray(ref.expect[k])).max() for k in range(2))
    print(f"non-Hermitian rho0: matrix form vs superoperator {err:.2e}", flush=True)
    assert err < 1e-5
    # a Hermitian state on the same solver switches the fast path back on
    solver.run(qa.basis(N, 0), tl, e_ops=e)
    assert solver.rhs.assume_hermitian_state is True
    assert solver._integrator._engine.hermitian is True


@pytest.mark.parametrize("N", [128, 256, 512])
def test_dense_shapes_agree_and_autotuner_picks_one(N):
    """Every tile shape the autotuner may pick gives the same solve to the
    FP32 products' rounding; invalid shapes are refused; the tuned shape is
    one of the candidates."""
    from qutip_amd import autotune
    print(f"dense shapes N={N}", flush=True)
    H, cs, Hnh, rho0 = _system(N, 3 + N, nc=1)
    tl = np.linspace(0, 0.3, 4)
    ops = [H]
    ref = None
    for herm in (True, False):
        eng = RKEngine(None, "vern7", 1e-8, 1e-10, dense=(Hnh, cs), dense_precision="fp32")
        eng.set_hermitian(herm, tune=False)
        default = eng.dense_shape
        for shape in autotune.dense_candidates(N, True, herm, default):
            eng.dense_shape = shape
            assert eng.dense_shape == shape
            eng.set_functionals(*_trace_functionals(ops, N))
            eng.set_state(tl[0], rho0.ravel(order="F"))
            vals, _, _ = eng.run(tl[1:])
            if ref is None:
                ref = vals
            diff = np.abs(vals - ref).max() / max(1.0, np.abs(ref).max())
            print(f"  herm={herm} shape={shape}: rel diff {diff:.1e}", flush=True)
            assert diff < 2e-6
        with pytest.raises(ValueError):  # an argument error of the C ABI
            eng.dense_shape = (0, 3, 1)
        tuned = autotune.dense_shape_for(eng)
        assert tuned in autotune.dense_candidates(N, True, herm, default)
