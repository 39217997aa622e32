"""The fused engine's output launches.

Each functional is summed in chunks of 16 entries, one wavefront each, for
every pending output time; with one chunk per functional the sums are the
values, otherwise the last wavefront to arrive adds the chunks' partial
sums, and it completes the output.  A run that stores no states leaves rk_store_kernel out of
its iterations and writes the final state once at the end.

* the one-chunk form against the chunked form on the same solve: the same
  step sequence, and <O>(t) equal to the summation order's rounding (1e-13
  relative), for linear (density matrix) and quadratic (ket) functionals,
  with one to several output times per step;
* the final state of a run without stored states is bit for bit the one of
  the same run with them (rk_store_kernel's), and both equal the last stored
  state; on a single-step run (step=True) as well."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import qutip_amd as qa  # noqa: E402
from qutip_amd.solver.hip_rk import RKEngine  # noqa: E402
from oracle import cbind, models  # noqa: E402


def _padded(fun, n, extra, quad):
    """fun plus one functional of `extra` entries (more than 16: every
    functional is then summed from partial sums)."""
    kinds, fptr, rows, cols, w = fun
    rng = np.random.default_rng(5)
    ecols = rng.integers(0, n, extra).astype(np.int32)
    erows = rng.integers(0, n, extra).astype(np.int32) if quad else np.zeros(extra, np.int32)
    ew = rng.standard_normal(extra) + 1j * rng.standard_normal(extra)
    return (np.append(kinds, 1 if quad else 0).astype(np.int32),
            np.append(fptr, fptr[-1] + extra).astype(np.int32),
            np.concatenate([rows, erows]).astype(np.int32),
            np.concatenate([cols, ecols]).astype(np.int32), np.concatenate([w, ew]))


def _run(eng, fun, y0, tl):
    eng.set_functionals(*fun)
    eng.set_state(tl[0], y0)
    vals, st, _ = eng.run(tl[1:])
    assert st >= 0
    return vals, eng.stats()


def _linear(ops, d):
    fptr, fidx, fw = cbind.functionals_from_ops(ops, d)
    nf = len(fptr) - 1
    return (np.zeros(nf, np.int32), fptr.astype(np.int32), np.zeros(len(fidx), np.int32),
            fidx.astype(np.int32), fw.view(np.complex128))


def _quadratic(ops):
    kinds, fptr, rows, cols, ws = [], [0], [], [], []
    for O in ops:
        O = O.tocoo()
        kinds.append(1)
        rows.append(O.row.astype(np.int32))
        cols.append(O.col.astype(np.int32))
        ws.append(O.data.astype(np.complex128))
        fptr.append(fptr[-1] + O.nnz)
    return (np.array(kinds, np.int32), np.array(fptr, np.int32), np.concatenate(rows),
            np.concatenate(cols), np.concatenate(ws))


@pytest.mark.parametrize("npts", [11, 201])
def test_one_chunk_output_matches_chunked_output_density_matrix(npts):
    print(f"output forms: TFIM L=4 Liouvillian, {npts} output times", flush=True)
    m = models.ising(4, 0.1)
    Lm = models.liouvillian(m["H"], m["c_ops"]).tocsr()
    d = m["H"].shape[0]
    rho0 = models.vec(models.ket2dm(m["psi0"]))
    tl = np.linspace(0.0, 2.0, npts)  # 201 points: several output times per step
    fun = _linear(m["e_ops"], d)
    assert np.diff(fun[1]).max() <= 16
    eng = RKEngine(Lm, "vern7", 1e-8, 1e-10, lanes=1)
    one, s1 = _run(eng, fun, rho0, tl)
    grid, s2 = _run(eng, _padded(fun, Lm.shape[0], 1100, False), rho0, tl)
    assert s1["attempts"] == s2["attempts"] and s1["dense_preps"] == s2["dense_preps"]
    nf = len(fun[0])
    diff = np.abs(one - grid[:, :nf]).max() / np.abs(grid[:, :nf]).max()
    print(f"  attempts {s1['attempts']}, max rel |one-chunk - chunked| {diff:.1e}", flush=True)
    assert diff < 1e-13


def test_one_chunk_output_matches_chunked_output_ket():
    print("output forms: TFIM L=4 ket, quadratic functionals", flush=True)
    m = models.ising(4, 0.1)
    H = (-1j * m["H"]).tocsr()
    psi0 = m["psi0"].astype(np.complex128)
    tl = np.linspace(0.0, 2.0, 41)
    fun = _quadratic(m["e_ops"])
    assert np.diff(fun[1]).max() <= 16
    eng = RKEngine(H, "vern7", 1e-8, 1e-10, lanes=1)
    one, s1 = _run(eng, fun, psi0, tl)
    grid, s2 = _run(eng, _padded(fun, H.shape[0], 1100, True), psi0, tl)
    assert s1["attempts"] == s2["attempts"]
    nf = len(fun[0])
    diff = np.abs(one - grid[:, :nf]).max() / np.abs(grid[:, :nf]).max()
    print(f"  attempts {s1['attempts']}, max rel |one-chunk - chunked| {diff:.1e}", flush=True)
    assert diff < 1e-13


def test_final_state_without_stored_states():
    print("final state: rk_final_state_kernel vs rk_store_kernel", flush=True)
    m = models.ising(5, 0.1)
    Lm = models.liouvillian(m["H"], m["c_ops"]).tocsr()
    d = m["H"].shape[0]
    rho0 = models.vec(models.ket2dm(m["psi0"]))
    fun = _linear(m["e_ops"], d)
    eng = RKEngine(Lm, "vern7", 1e-8, 1e-10, lanes=1)
    eng.set_functionals(*fun)
    for tl in (np.linspace(0.0, 1.3, 14), np.array([0.0, 0.7])):
        eng.set_state(tl[0], rho0)
        eng.run(tl[1:])
        t_a, ya = eng.get_state()
        ya = ya.to_array().ravel()
        buf = qa.HipDense((eng.npad, len(tl) - 1))
        eng.set_state(tl[0], rho0)
        eng.run(tl[1:], states_out=buf.ptr)
        t_b, yb = eng.get_state()
        yb = yb.to_array().ravel()
        last = buf.to_array()[:Lm.shape[0], -1]
        assert t_a == t_b == tl[-1]
        assert np.array_equal(ya, yb)
        assert np.array_equal(yb, last)
    # single steps: the final state is the step's front
    eng.set_state(0.0, rho0)
    fronts = []
    for _ in range(3):
        eng.run([1.0], single_step=True)
        t, y = eng.get_state()
        fronts.append((t, y.to_array()))
    assert 0.0 < fronts[0][0] < fronts[1][0] < fronts[2][0] <= 1.0
    eng.set_state(0.0, rho0)
    for t_ref, y_ref in fronts:
        buf = qa.HipDense((eng.npad, 1))
        eng.run([1.0], single_step=True, states_out=buf.ptr)
        t, y = eng.get_state()
        assert t == t_ref and np.array_equal(y.to_array(), y_ref)
