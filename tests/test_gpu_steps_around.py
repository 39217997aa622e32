"""The steps around the path on the GPU: steadystate by repeated
propagation (the fused mesolve engine) against the direct method, and the
merge of two mcsolve results run on disjoint seed sets (how a sharded
ensemble is combined) against one run over all the seeds."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import qutip_amd as qa  # noqa: E402
from tests import systems  # noqa: E402


def test_steadystate_propagator_vs_direct():
    N, k, nth = 12, 0.5, 0.3
    a = qa.destroy(N)
    H = 0.8 * a.dag() @ a + 0.2 * (a + a.dag())
    c_ops = [np.sqrt(k * (1 + nth)) * a, np.sqrt(k * nth) * a.dag()]
    print("steadystate propagator vs direct", flush=True)
    direct = qa.steadystate(H, c_ops)
    prop = qa.steadystate(H, c_ops, method="propagator", rho=qa.ket2dm(qa.basis(N, 3)),
                          propagator_tol=1e-12)
    err = np.abs(prop.full() - direct.full()).max()
    print(f"  max |rho_prop - rho_direct| {err:.2e}", flush=True)
    assert err < 1e-6


def test_mcsolve_merge_of_seed_halves():
    s = systems.ising(4, 0.2, tf=4.0, npts=9)
    solver = qa.MCSolver(s["H"], s["c_ops"])
    n = 256
    print(f"mcsolve merge of two halves of {n} seeds", flush=True)
    full = solver.run(s["psi0"], s["tlist"], n, e_ops=s["e_ops"], seeds=list(range(n)))
    a = solver.run(s["psi0"], s["tlist"], n // 2, e_ops=s["e_ops"], seeds=list(range(n // 2)))
    b = solver.run(s["psi0"], s["tlist"], n // 2, e_ops=s["e_ops"],
                   seeds=list(range(n // 2, n)))
    m = a + b
    assert m.num_trajectories == n
    for k in range(2):
        d_avg = np.abs(np.asarray(m.average_expect[k]) - np.asarray(full.average_expect[k])).max()
        d_std = np.abs(np.asarray(m.std_expect[k]) - np.asarray(full.std_expect[k])).max()
        print(f"  e_op {k}: |avg diff| {d_avg:.1e}, |std diff| {d_std:.1e}", flush=True)
        assert d_avg < 1e-13 and d_std < 1e-12
    assert list(m.num_jumps) == list(full.num_jumps)
