"""hipGraph replay of the fused engine against direct launches.

The engine captures one loop iteration into a hipGraph and replays it; with
QUTIP_AMD_GRAPH=0 (read when an engine is created) it enqueues the same
kernels directly.  Each engine kind solves several times on the graph, with
other device work and synchronisations between the solves, and every solve
must give exactly the bits of the direct-launch engine: the same kernels in
the same order on the same data.  Bar: bit-identical."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import qutip_amd as qa  # noqa: E402
from qutip_amd import _lib  # noqa: E402
from qutip_amd.solver.hip_rk import RKEngine  # noqa: E402
from oracle import cbind, models  # noqa: E402


def _solve(eng, y0, tl, fun):
    fptr, fidx, fw = fun
    eng.set_functionals(np.zeros(len(fptr) - 1, np.int32), fptr.astype(np.int32),
                        np.zeros(len(fidx), np.int32), fidx.astype(np.int32),
                        fw.view(np.complex128))
    eng.set_state(tl[0], y0)
    vals, st, _ = eng.run(tl[1:])
    assert st >= 0
    _, y = eng.get_state()
    return vals, y.to_array().ravel(), eng.stats()["attempts"]


def _engines(monkeypatch, make):
    monkeypatch.setenv("QUTIP_AMD_GRAPH", "0")
    eager = make()
    monkeypatch.delenv("QUTIP_AMD_GRAPH")
    return eager, make()


def _check_replays(eager, graph, y0, tl, fun, other_work):
    ref = _solve(eager, y0, tl, fun)
    for rep in range(3):
        got = _solve(graph, y0, tl, fun)
        assert got[2] == ref[2], rep
        assert np.array_equal(got[0], ref[0]), rep
        assert np.array_equal(got[1], ref[1]), rep
        other_work()
        _lib.call("qa_device_sync")


def test_sparse_coded_engine_graph_replays(monkeypatch):
    print("graph replays: coded sparse engine (TFIM L=5 Liouvillian)", flush=True)
    m = models.ising(5, 0.1)
    Lm = models.liouvillian(m["H"], m["c_ops"]).tocsr()
    d = m["H"].shape[0]
    rho0 = models.vec(models.ket2dm(m["psi0"]))
    tl = np.linspace(0.0, 1.0, 11)
    fun = cbind.functionals_from_ops(m["e_ops"], d)
    eager, graph = _engines(monkeypatch, lambda: RKEngine(Lm, "vern7", 1e-8, 1e-10, lanes=1))
    assert graph.coded
    x = qa.HipDense.from_array(np.random.default_rng(0).standard_normal(Lm.shape[0]) + 0j)
    _check_replays(eager, graph, rho0, tl, fun, lambda: eager.apply(x))


@pytest.mark.parametrize("prec", ["fp64_banded", "fp32"])
def test_matrix_form_engine_graph_replays(monkeypatch, prec):
    print(f"graph replays: matrix-form engine {prec}", flush=True)
    N = 32
    a = np.diag(np.sqrt(np.arange(1, N)), 1).astype(complex)
    H = 0.5 * a.conj().T @ a + 0.05 * a.conj().T @ a.conj().T @ a @ a + 0.3 * (a + a.conj().T)
    c = np.sqrt(0.5) * a
    Hnh = H - 0.5j * c.conj().T @ c
    rho0 = np.zeros((N, N), complex)
    rho0[0, 0] = 1.0
    y0 = rho0.ravel(order="F")
    tl = np.linspace(0.0, 1.0, 6)
    # functionals: tr(rho) and <n> on the column-major vec
    fptr = np.array([0, N, 2 * N], np.int32)
    fidx = np.concatenate([np.arange(N) * (N + 1)] * 2).astype(np.int32)
    fw = np.concatenate([np.ones(N), np.arange(N)]).astype(np.complex128).view(np.float64)
    eager, graph = _engines(
        monkeypatch, lambda: RKEngine(None, "vern7", 1e-8, 1e-10, dense=(Hnh, [c]),
                                      dense_precision=prec))
    assert graph.dense_precision == prec
    x = qa.HipDense.from_array(np.eye(N, dtype=complex).ravel(order="F"))
    _check_replays(eager, graph, y0, tl, (fptr, fidx, fw), lambda: eager.apply(x))
