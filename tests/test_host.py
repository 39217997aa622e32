"""Host-side logic of the product (CPU only): SELL-64 layout, output
functionals, the object model against the oracle's independent
construction, option handling."""
import numpy as np
import pytest
import scipy.sparse as sp
from numpy.testing import assert_allclose

import qutip_amd as qa
from qutip_amd import data as qdata
from oracle import models
from tests import systems


def _sell_matvec(sptr, cols, vals, n, x, lanes):
    """numpy emulation of the SELL-C kernel: lane l = lanes r + q of slice s
    sums entries q, q + lanes, ... of row C s + r."""
    ns = len(sptr) - 1
    C = 64 // lanes
    out = np.zeros(ns * C, dtype=np.complex128)
    for s in range(ns):
        steps = (sptr[s + 1] - sptr[s]) // 64
        blk_c = cols[sptr[s]:sptr[s + 1]].reshape(steps, 64)
        blk_v = vals[sptr[s]:sptr[s + 1]].reshape(steps, 64)
        lane_sum = (blk_v * x[blk_c]).sum(axis=0)
        out[s * C:(s + 1) * C] = lane_sum.reshape(C, lanes).sum(axis=1)
    return out[:n]


@pytest.mark.parametrize("lanes", [1, 2, 4, 8])
@pytest.mark.parametrize("n,density", [(1, 1.0), (63, 0.1), (64, 0.05), (200, 0.03), (130, 0.0)])
def test_sell_layout_reproduces_csr(n, density, lanes):
    rng = np.random.default_rng(n)
    A = sp.random(n, n, density=density, random_state=rng, format="csr") * (1 + 0.5j)
    sptr, cols, vals, npad = qdata.sell_arrays(A, lanes)
    assert npad % (64 // lanes) == 0 and npad >= n
    assert cols.min() >= 0 and cols.max() < n
    assert sptr[-1] % 64 == 0
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert_allclose(_sell_matvec(sptr, cols, vals, n, x, lanes), A @ x, atol=1e-13)


def test_sell_padding_overhead_on_ising():
    m = models.ising(8, 0.1)
    Lm = models.liouvillian(m["H"], m["c_ops"])
    sptr, _, _, _ = qdata.sell_arrays(Lm)
    assert sptr[-1] / Lm.nnz < 1.09  # 7.9 % padding


def test_product_liouvillian_equals_oracle():
    s = systems.jaynes_cummings()
    Lp = qa.liouvillian(s["H"], s["c_ops"]).data
    m = models.jaynes_cummings()
    Lo = models.liouvillian(m["H"], m["c_ops"])
    assert abs(Lp - Lo).max() < 1e-13
    s = systems.ising(4, 0.1)
    m = models.ising(4, 0.1)
    assert abs(qa.liouvillian(s["H"], s["c_ops"]).data - models.liouvillian(m["H"], m["c_ops"])).max() < 1e-13


def test_functionals_give_expectations():
    """The linear functionals of a vectorised rho reproduce tr(O rho)."""
    s = systems.jaynes_cummings()
    solver = qa.MESolver.__new__(qa.MESolver)
    solver._state_is_ket = False
    solver._state_shape = (20, 20)
    kinds, fptr, rows, cols, w = qa.Solver._functionals(solver, s["e_ops"])
    rng = np.random.default_rng(1)
    r = rng.standard_normal((20, 20)) + 1j * rng.standard_normal((20, 20))
    rho = r @ r.conj().T
    v = rho.ravel(order="F")
    for f, O in enumerate(s["e_ops"] + [qa.qeye(20)]):
        val = np.sum(w[fptr[f]:fptr[f + 1]] * v[cols[fptr[f]:fptr[f + 1]]])
        assert val == pytest.approx(models.expect_dm(sp.csr_matrix(O.data), rho), rel=1e-12)


def test_functionals_quadratic_ket():
    s = systems.ising(4, 0.1)
    solver = qa.SESolver.__new__(qa.SESolver)
    solver._state_is_ket = True
    solver._state_shape = (16, 1)
    kinds, fptr, rows, cols, w = qa.Solver._functionals(solver, s["e_ops"])
    assert list(kinds) == [1, 1, 1]
    rng = np.random.default_rng(2)
    psi = rng.standard_normal(16) + 1j * rng.standard_normal(16)
    for f, O in enumerate(s["e_ops"]):
        sl = slice(fptr[f], fptr[f + 1])
        val = np.sum(w[sl] * psi[rows[sl]].conj() * psi[cols[sl]])
        assert val == pytest.approx(models.expect_ket(sp.csr_matrix(O.data), psi), rel=1e-12)


def test_unknown_option_raises_keyerror():
    s = systems.jaynes_cummings()
    with pytest.raises(KeyError):
        qa.MESolver(s["H"], s["c_ops"], options={"not_an_option": 1})


def test_integrator_registry():
    assert qa.MESolver.avail_integrators()["hip_rk"] is qa.IntegratorHipRK
    assert qa.SESolver.avail_integrators()["hip_rk"] is qa.IntegratorHipRK
    assert qa.IntegratorHipRK.integrator_options["rtol"] == 1e-6


def test_device_ops_fail_loudly_without_gpu():
    import torch  # noqa: F401  (only to make sure no GPU is visible here)
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(qa._lib.QutipAmdUnavailable):
        qa.HipDense.from_array(np.ones(4))


def test_qutip_plugin_imports_without_qutip():
    from qutip_amd import qutip_plugin
    assert qutip_plugin.available() in (True, False)
    if not qutip_plugin.available():
        with pytest.raises(ImportError):
            qutip_plugin.register()


@pytest.mark.parametrize("d,density", [(1, 1.0), (7, 0.3), (64, 0.05), (130, 0.0), (33, 1.0)])
def test_ell_layout_reproduces_csr(d, density):
    """The trajectory kernel's column-major ELL: entry u of row r at u d + r,
    width a multiple of 4, padding with value 0 and an in-range column."""
    from qutip_amd.solver.mcsolve import _ell_host
    rng = np.random.default_rng(d)
    A = sp.random(d, d, density=density, random_state=rng, format="csr") * (1 - 0.25j)
    W, col, val = _ell_host(A, d, False)
    assert W % 4 == 0 and W >= np.diff(A.indptr).max()
    assert col.shape == (max(W, 1), d) and col.min() >= 0 and col.max() < d
    x = rng.standard_normal(d) + 1j * rng.standard_normal(d)
    y = (val[:W] * x[col[:W]]).sum(axis=0) if W else np.zeros(d, complex)
    assert_allclose(y, A @ x, atol=1e-13)
    _, _, val32 = _ell_host(A, d, True)
    assert val32.dtype == np.complex64


def _sell_coded_matvec(sptr, codes, vtab, otab, vbits, shape, x, lanes):
    """numpy emulation of the coded SELL-C kernel: word w of lane l of slice
    s is the uint64 at 8 (sptr[s] + 64 w + l) bytes holding the lane's steps
    4w .. 4w + 3 in its 16-bit quarters, low first; a code's low vbits bits
    pick the value, the rest the column offset from the row (clamped)."""
    n, ncols = shape
    ns = len(sptr) - 1
    C = 64 // lanes
    out = np.zeros(ns * C, dtype=np.complex128)
    for s in range(ns):
        words = (sptr[s + 1] - sptr[s]) // 64
        c = codes[4 * sptr[s]:4 * sptr[s + 1]].reshape(words, 64, 4).astype(np.int64)
        rows = s * C + np.arange(64) // lanes
        cols = np.clip(rows[None, :, None] + otab[c >> vbits], 0, ncols - 1)
        lane_sum = (vtab[c & ((1 << vbits) - 1)] * x[cols]).sum(axis=(0, 2))
        out[s * C:(s + 1) * C] = lane_sum.reshape(C, lanes).sum(axis=1)
    return out[:n]


@pytest.mark.parametrize("lanes", [1, 2, 4, 8])
def test_coded_layout_reproduces_csr(lanes):
    """Coded SELL-C of the dissipative Ising Liouvillian (L=5) and of a
    ragged banded matrix (n not a multiple of 64, explicit zeros): the
    kernel's reading of the codes gives A x, and slices are whole words of
    the 64 lanes."""
    m = models.ising(5, 0.1)
    Lm = models.liouvillian(m["H"], m["c_ops"]).tocsr()
    rng = np.random.default_rng(lanes)
    n = 1000
    band = sp.diags([rng.choice([0.5, -1j, 2.0], n - k) for k in (0, 1, 7, 130)],
                    [0, 1, -7, 130], shape=(n, n), format="csr")
    band.data[::17] = 0.0  # explicit zeros are entries like any other
    for A, uniform in ((Lm, "auto"), (Lm, True), (Lm, False), (band, "auto"), (band, True)):
        arrs = qdata.sell_coded_arrays(A, lanes, uniform=uniform)
        assert arrs is not None
        sptr, codes, vtab, otab, vbits, npad = arrs
        assert np.all(np.diff(sptr) % 64 == 0)
        assert len(vtab) == 1 << vbits and len(otab) == 1 << (16 - vbits)
        if uniform is True:
            assert len(set(np.diff(sptr).tolist())) == 1
        x = rng.standard_normal(A.shape[1]) + 1j * rng.standard_normal(A.shape[1])
        assert_allclose(_sell_coded_matvec(sptr, codes, vtab, otab, vbits, A.shape, x, lanes),
                        A @ x, atol=1e-12)
    # the same value bits as the plain form: its values appear in the table
    sptr, codes, vtab, otab, vbits, _ = qdata.sell_coded_arrays(Lm, lanes)
    assert set(Lm.data.view(np.uint64).reshape(-1, 2).tobytes()[i:i + 16]
               for i in range(0, 16 * Lm.nnz, 16)) <= set(
        vtab.view(np.uint64).reshape(-1, 2).tobytes()[i:i + 16] for i in range(0, 16 * 256, 16))


def test_coded_layout_limits():
    """The 16-bit split: value bits b in 6..10 with <= 2^b values and
    <= 2^(16-b) column offsets, the smallest b that fits; no coded form when
    none fits.  -0.0 and 0.0 are distinct values (their products can differ
    in sign)."""
    assert qdata.code_split(64, 1024) == 6 and qdata.code_split(65, 1024) is None
    assert qdata.code_split(352, 41) == 9 and qdata.code_split(1024, 64) == 10
    assert qdata.code_split(1025, 1) is None
    rng = np.random.default_rng(0)
    A = sp.random(300, 300, density=0.05, random_state=rng, format="csr")
    assert qdata.sell_coded_arrays(A) is None  # ~4500 distinct values
    # row r at column r + r // 2 (mod n): 300 distinct column offsets
    n = 600
    cols = (np.arange(n) + np.arange(n) // 2) % n
    assert len(np.unique(cols - np.arange(n))) == 300
    P = sp.csr_matrix((np.ones(n), (np.arange(n), cols)), shape=(n, n))
    arrs = qdata.sell_coded_arrays(P)
    assert arrs is not None and arrs[4] == 6  # 2 values (1 and the padding 0)
    P2 = sp.csr_matrix((np.arange(1.0, n + 1), (np.arange(n), cols)), shape=(n, n))
    assert qdata.sell_coded_arrays(P2) is None  # 601 values with 300 offsets
    Z = sp.csr_matrix((np.array([-0.0, 1.0]), np.array([1, 0]), np.array([0, 1, 2])), shape=(2, 2))
    sptr, codes, vtab, otab, vbits, _ = qdata.sell_coded_arrays(Z)
    signs = {np.signbit(v.real) for v in vtab[codes[:4] & ((1 << vbits) - 1)]}
    assert signs == {True, False}
    with pytest.raises(ValueError, match="code split"):
        qdata.SellMatrix(A, coded=True)


def test_banded_operators_reproduce_the_matrices():
    """The diagonals handed to the banded matrix-form kernel: row-indexed
    (A[r, r + o] at [d, r]), one operator index per diagonal, and None past
    96 products of diagonals."""
    from qutip_amd.solver.hip_rk import banded_operators
    rng = np.random.default_rng(5)
    N = 21
    H = np.diag(rng.standard_normal(N)) + np.diag(rng.standard_normal(N - 2) + 1j, 2)
    c = np.diag(np.sqrt(np.arange(1, N)), 1) + np.diag(np.ones(N - 5), -5)
    dop, doff, dval = banded_operators(H, [c])
    for k, A in enumerate((H, c)):
        R = np.zeros((N, N), complex)
        for d in np.nonzero(dop == k)[0]:
            o = doff[d]
            rr = np.arange(max(0, -o), min(N, N - o))
            R[rr, rr + o] = dval[d, rr]
        assert np.array_equal(R, A)
    dense = rng.standard_normal((N, N))
    assert banded_operators(H, [dense]) is None  # 41 diagonals: 1681 products


def test_server_fails_a_request_above_the_budget():
    """A request that can never fit the HBM budget fails at once instead of
    waiting forever (no device work is started)."""
    from qutip_amd.serve import SolveRequest, SolveServer
    req = SolveRequest("se", qa.sigmax(), qa.basis(2, 0), [0.0, 1.0], e_ops=[qa.sigmaz()])
    done = SolveServer(hbm_budget=1).run([req], realtime=False)
    assert done[0].error and "HBM budget" in done[0].error
    assert done[0].latency() is not None and np.isfinite(done[0].latency())


def test_shortest_work_first_estimates():
    from qutip_amd.serve import SolveServer
    reqs = systems.serving_trace(30)
    w = {r.id: r.work_estimate() for r in reqs}
    me512 = [r for r in reqs if r.kind == "me" and r.d == 512]
    se16 = [r for r in reqs if r.kind == "se" and r.d == 16]
    if me512 and se16:
        assert w[me512[0].id] > w[se16[0].id]
    with pytest.raises(ValueError):
        SolveServer(policy="lifo")


def test_solver_functionals_kept_per_e_ops_list():
    """Solver._functionals builds the device functionals of an e_ops list
    once: the same operators (same objects, same data) give back the kept
    arrays; a reordered list, an operator whose data was replaced or another
    state shape rebuild them, and every result equals a fresh build."""
    from qutip_amd.solver.solver_base import Solver

    def fresh(s, ops):
        t = Solver.__new__(Solver)
        t._state_is_ket, t._state_shape = s._state_is_ket, s._state_shape
        return t._build_functionals(ops)

    def same(a, b):
        return all(np.array_equal(x, y) for x, y in zip(a, b))

    s = Solver.__new__(Solver)
    s._state_is_ket, s._state_shape = False, (6, 6)
    A, B = qa.num(6), qa.destroy(6) + qa.create(6)
    f1 = s._functionals([A, B])
    assert same(f1, fresh(s, [A, B]))
    assert s._functionals([A, B]) is f1  # kept
    f2 = s._functionals([B, A])
    assert f2 is not f1 and same(f2, fresh(s, [B, A]))
    B.data = (2.0 * B).data  # new data object on the same Qobj
    f3 = s._functionals([B, A])
    assert f3 is not f2 and same(f3, fresh(s, [B, A]))
    assert not same(f3, f2)
    s._state_is_ket, s._state_shape = True, (6, 1)  # a ket: quadratic forms
    f4 = s._functionals([B, A])
    assert same(f4, fresh(s, [B, A])) and (f4[0] == 1).all()


@pytest.mark.parametrize("N,density", [(40, 0.1), (64, 0.3), (24, 0.7)])
def test_rand_herm_fill_distribution(N, density):
    """rand_herm follows the "fill" distribution of random_objects.py:198-327:
    Hermitian, entries of (X + X^dag)/2 with X uniform in the [-1,1]^2 box,
    and the reference's entry count (sparse: that many distinct positions;
    dense: that many symmetric pairs removed)."""
    H = qa.rand_herm(N, density=density, seed=3).full()
    assert np.array_equal(H, H.conj().T)
    assert np.abs(H.real).max() <= 1.0 and np.abs(H.imag).max() <= 1.0
    nz = np.count_nonzero(H)
    if density < 0.5:
        num = int(max((N * N - 0.666 * N) * (1 - (1 - density) ** 0.5) + 0.666 * N * density, 1))
        # each drawn position and its mirror; a diagonal or mirrored pair
        # coincides with itself
        assert num <= nz <= 2 * num
    else:
        num = int(max(N * (N - 0.666) * (1 - density ** 0.5) + 0.666 * N * (1 - density), 1))
        assert N * N - 2 * num <= nz <= N * N - num
    # the same seed gives the same matrix; another seed another one
    assert np.array_equal(H, qa.rand_herm(N, density=density, seed=3).full())
    assert not np.array_equal(H, qa.rand_herm(N, density=density, seed=4).full())


def test_rand_dm_ginibre():
    """rand_dm is a full-rank Ginibre density matrix (random_objects.py:
    612-643): Hermitian, unit trace, positive definite; X has standard
    complex normal entries, so E tr(X X^dag) = 2 N^2."""
    N = 48
    rho = qa.rand_dm(N, seed=9).full()
    assert np.abs(rho - rho.conj().T).max() < 1e-15
    assert abs(np.trace(rho) - 1) < 1e-12
    assert np.linalg.eigvalsh(rho).min() > 0
    g = np.random.default_rng(9).normal(size=(N, N, 2))
    X = g[..., 0] + 1j * g[..., 1]
    assert abs(np.sum(np.abs(X) ** 2) / (2 * N * N) - 1) < 0.05


def test_qobjevo_list_form():
    """QobjEvo([H0, [H1, c]]) evaluates to H0 + c(t) H1 (qobjevo.pyx:437-480);
    algebra keeps the coefficients; dag conjugates them."""
    sx, sz = qa.sigmax(), qa.sigmaz()
    H = qa.QobjEvo([sz, [sx, "cos(w * t)"], [sz, np.array([0.0, 1.0, 2.0])]],
                   args={"w": 2.0}, tlist=np.array([0.0, 1.0, 2.0]))
    assert not H.isconstant and len(H.elements) == 2
    for t in (0.0, 0.4, 1.7):
        want = sz.full() * (1 + t) + np.cos(2 * t) * sx.full()
        assert np.allclose(H(t).full(), want)
    G = (2.0 * H - sz).dag()
    assert np.allclose(G(0.3).full(), (2 * H(0.3) - sz).full().conj().T)
    H.arguments(w=3.0)
    assert np.allclose(H(0.5).full(), sz.full() * 1.5 + np.cos(1.5) * sx.full())
    assert isinstance(H.to_list()[1], list)


def test_liouvillian_of_time_dependent_h():
    """liouvillian of a list-form H: the constant Liouvillian plus -i(spre -
    spost) of each element with its coefficient; at any t it equals the
    oracle's Liouvillian of H(t)."""
    a = qa.destroy(4)
    H = qa.QobjEvo([a.dag() @ a, [a + a.dag(), "sin(t)"]])
    L = qa.liouvillian(H, [np.sqrt(0.2) * a])
    for t in (0.0, 0.9):
        want = models.liouvillian(H(t).to_csr(), [np.sqrt(0.2) * a.to_csr()])
        assert abs(L(t).to_csr() - want).max() < 1e-14


def test_time_dependent_routing_on_the_host():
    """Coefficients with device programs give the fused time-dependent
    engine its operator; a Python function coefficient leaves it to the
    plain form."""
    from qutip_amd.solver.hip_rk import time_dependent_operator
    sx, sz = qa.sigmax(), qa.sigmaz()
    rhs = -1j * qa.QobjEvo([sz, [sx, "cos(t)"]])
    A0, parts = time_dependent_operator(rhs.matmul_data)
    assert abs(A0 - (-1j * sz.to_csr())).max() == 0 and len(parts) == 1
    assert parts[0][1].evaluate(0.3) == pytest.approx(np.cos(0.3))
    rhs_f = -1j * qa.QobjEvo([sz, [sx, lambda t: np.cos(t)]])
    assert time_dependent_operator(rhs_f.matmul_data) is None
    assert time_dependent_operator(qa.QobjEvo(sz).matmul_data) is None


def test_stacked_elements_code_ranges():
    """The stacked operator keeps every element's entries apart and its
    coded value table is split by element (ranges in order)."""
    rng = np.random.default_rng(1)
    mats = [sp.random(40, 40, density=0.1, random_state=rng, format="csr") * (k + 1)
            for k in range(3)]
    A, elem = qdata.stacked_elements(mats)
    assert A.nnz == sum(m.nnz for m in mats)
    assert abs(A - sum(mats)).max() < 1e-15
    arrs = qdata.sell_coded_arrays(A, 1, elem=elem)
    vstart = arrs[-1]
    assert list(vstart) == sorted(vstart) and vstart[0] == 0 and len(vstart) == 4
    vtab = arrs[2]
    for k, m in enumerate(mats):
        vals = set(np.round(vtab[vstart[k]:vstart[k + 1]], 12)) - {0}
        assert vals == set(np.round(m.data, 12)) - {0}
