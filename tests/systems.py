"""Benchmark systems built with the product's own Qobj API (the way the
reference's tests build theirs with qutip), shared by the GPU tests and
bench.py.  The oracle builds the same systems independently in
oracle/models.py."""
import numpy as np

import qutip_amd as qa


def jaynes_cummings(N=10, g=0.05 * 2 * np.pi):
    a = qa.tensor(qa.destroy(N), qa.qeye(2))
    sm = qa.tensor(qa.qeye(N), qa.sigmam())
    sz = qa.tensor(qa.qeye(N), qa.sigmaz())
    wc = wq = 2 * np.pi
    H = wc * a.dag() @ a + wq * 0.5 * sz + g * (a @ sm.dag() + a.dag() @ sm)
    c_ops = [np.sqrt(0.005) * a, np.sqrt(0.05) * sm]
    psi0 = qa.tensor(qa.basis(N, 0), qa.basis(2, 0))
    return dict(H=H, c_ops=c_ops, psi0=psi0, e_ops=[a.dag() @ a, sz],
                tlist=np.linspace(0, 25, 100))


def site(op, i, L):
    return qa.tensor([op if j == i else qa.qeye(2) for j in range(L)])


def ising(L=8, gamma=0.1, tf=10.0, npts=200, seed=0):
    h = np.random.default_rng(seed).uniform(0.5, 1.5, L)
    H = 0
    for i in range(L - 1):
        H = H + site(qa.sigmaz(), i, L) @ site(qa.sigmaz(), i + 1, L)
    for i in range(L):
        H = H + h[i] * site(qa.sigmax(), i, L)
    c_ops = [np.sqrt(gamma) * site(qa.sigmam(), i, L) for i in range(L)]
    psi0 = qa.tensor([qa.basis(2, 0)] * L)
    mz = site(qa.sigmaz(), 0, L)
    for i in range(1, L):
        mz = mz + site(qa.sigmaz(), i, L)
    mz = mz * (1.0 / L)
    zz = site(qa.sigmaz(), 0, L) @ site(qa.sigmaz(), 1, L)
    return dict(H=H, c_ops=c_ops, psi0=psi0, e_ops=[mz, zz],
                tlist=np.linspace(0, tf, npts), h=h)


def kerr(N=256, delta=0.5, K=0.1, F=2.0, kappa=1.0):
    a = qa.destroy(N)
    H = delta * a.dag() @ a + 0.5 * K * a.dag() @ a.dag() @ a @ a + F * (a + a.dag())
    parity = qa.Qobj(np.diag((-1.0) ** np.arange(N)), dims=[[N], [N]])
    return dict(H=H, c_ops=[np.sqrt(kappa) * a], psi0=qa.basis(N, 0),
                e_ops=[a.dag() @ a, parity], tlist=np.linspace(0, 30, 150))


SERVING_OPTIONS = {"rtol": 1e-8, "atol": 1e-10}


def serving_trace(n, rate=200.0, seed=7, dims=(16, 32, 64, 128, 256, 512), tf=10.0, npts=50,
                  kinds=("se", "me", "mc"), kind_p=(0.4, 0.4, 0.2), ntraj=(64, 512),
                  options=None):
    """The serving workload: Poisson arrivals (rate per second), d uniform in
    ``dims``, H = rand_herm(d, density=0.1), 1-4 collapse operators
    sqrt(gamma) rand_dm(d) with gamma ~ U[0.01, 0.1], 40% sesolve / 40% mesolve
    / 20% mcsolve (ntraj ~ U{64..512}), t in [0, 10] with 50 outputs,
    psi0 = basis(d, 0), e_ops = {|0><0|, H}; every draw from default_rng(seed).
    Solver options: ``options`` (default rtol 1e-8, atol 1e-10, the
    tolerances the <O> error bar of 1e-6 is quoted at)."""
    options = dict(SERVING_OPTIONS if options is None else options)
    from qutip_amd.serve import SolveRequest
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        d = int(rng.choice(dims))
        kind = str(rng.choice(kinds, p=kind_p))
        nc = int(rng.integers(1, 5))
        gammas = rng.uniform(0.01, 0.1, nc)
        nt = int(rng.integers(ntraj[0], ntraj[1] + 1))
        s_op = int(rng.integers(2 ** 31))
        H = qa.rand_herm(d, density=0.1, seed=s_op)
        c_ops = [np.sqrt(g) * qa.rand_dm(d, seed=s_op + 1 + k) for k, g in enumerate(gammas)]
        proj = qa.basis(d, 0) @ qa.basis(d, 0).dag()
        reqs.append(SolveRequest(kind, H, qa.basis(d, 0), np.linspace(0, tf, npts),
                                 c_ops=c_ops if kind != "se" else [], e_ops=[proj, H],
                                 ntraj=nt if kind == "mc" else 0, seed=s_op, arrival=t, rid=i,
                                 options=options))
    return reqs
