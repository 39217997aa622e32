"""Benchmark systems and the Liouvillian, restated in scipy (test oracle).

Conventions restated from the reference:

* ``sigmaz = diag(1, -1)``, ``sigmax``, ``sigmam = [[0, 0], [1, 0]]``,
  ``sigmap = sigmam^T``: qutip/core/operators.py:353-445.  ``basis(2, 0)`` is
  spin-up / the excited qubit state (sigmaz = +1).
* ``destroy(N)``: off-diagonal sqrt(n), qutip/core/operators.py:468-505.
* tensor products in the order written: qutip/core/tensor.py (kron).
* column-stacking vectorisation vec(rho)[c * d + r] = rho[r, c]:
  qutip/core/superoperator.py:291-334.
* ``spre(A) = I (x) A``, ``spost(A) = A^T (x) I``: superoperator.py:379, 409.
* ``liouvillian(H, c_ops) = -i (spre(H) - spost(H))
  + sum_c [conj(c) (x) c - 1/2 I (x) c^dag c - 1/2 (c^dag c)^T (x) I]``:
  superoperator.py:123-134 (data branch), lindblad_dissipator :161-229.
* mcsolve effective Hamiltonian ``H_eff = -i H - 1/2 sum_c c^dag c``:
  qutip/solver/mcsolve.py:491-496.

The four systems are the benchmark configurations (sizes in the docstrings of
each builder).  The Ising field draw ``h = default_rng(0).uniform(0.5, 1.5, L)``
is this project's fixed choice for "h_i ~ U[0.5, 1.5] from seed".
"""
import numpy as np
import scipy.sparse as sp

CD = np.complex128


def sigmaz():
    return sp.csr_matrix(np.array([[1, 0], [0, -1]], dtype=CD))


def sigmax():
    return sp.csr_matrix(np.array([[0, 1], [1, 0]], dtype=CD))


def sigmam():
    return sp.csr_matrix(np.array([[0, 0], [1, 0]], dtype=CD))


def sigmap():
    return sp.csr_matrix(np.array([[0, 1], [0, 0]], dtype=CD))


def destroy(n):
    return sp.diags(np.sqrt(np.arange(1, n, dtype=np.float64)), 1, shape=(n, n),
                    dtype=CD, format="csr")


def qeye(n):
    return sp.identity(n, dtype=CD, format="csr")


def basis(n, k):
    v = np.zeros(n, dtype=CD)
    v[k] = 1.0
    return v


def kron_all(ops):
    out = sp.csr_matrix(np.ones((1, 1), dtype=CD))
    for o in ops:
        out = sp.kron(out, o, format="csr")
    return out


def site_op(op, i, L):
    return kron_all([op if j == i else qeye(2) for j in range(L)])


def dag(a):
    return a.conj().T.tocsr()


def liouvillian(H, c_ops):
    d = H.shape[0]
    eye = qeye(d)
    Lm = -1j * (sp.kron(eye, H) - sp.kron(H.T, eye))
    for c in c_ops:
        cdc = dag(c) @ c
        Lm = Lm + sp.kron(c.conj(), c) - 0.5 * sp.kron(eye, cdc) - 0.5 * sp.kron(cdc.T, eye)
    Lm = sp.csr_matrix(Lm, dtype=CD)
    Lm.sum_duplicates()
    Lm.eliminate_zeros()
    Lm.sort_indices()
    return Lm


def heff(H, c_ops):
    Hf = -1j * H
    for c in c_ops:
        Hf = Hf - 0.5 * (dag(c) @ c)
    Hf = sp.csr_matrix(Hf, dtype=CD)
    Hf.sum_duplicates()
    Hf.eliminate_zeros()
    Hf.sort_indices()
    return Hf


def matrix_form_rhs(H, c_ops):
    """The matrix-form Lindblad right-hand side on vec(rho), following
    LindbladMatrixForm.matmul_data with assume_hermitian_state
    (qutip/core/cy/lindblad_matrix_form.pyx:170-186): H_nh = H - i/2 sum
    c^dag c (:92-97), A = -i H_nh rho + 1/2 sum c rho c^dag, drho/dt =
    A + A^dag.  Dense numpy arrays; returns f(t, y) for oracle/rk.py."""
    H = np.asarray(H.toarray() if sp.issparse(H) else H, dtype=CD)
    cs = [np.asarray(c.toarray() if sp.issparse(c) else c, dtype=CD) for c in c_ops]
    Hnh = H.copy()
    for c in cs:
        Hnh = Hnh - 0.5j * (c.conj().T @ c)
    d = H.shape[0]

    def f(t, y):
        rho = y.reshape((d, d), order="F")
        A = -1j * (Hnh @ rho)
        for c in cs:
            A = A + 0.5 * ((c @ rho) @ c.conj().T)
        return (A + A.conj().T).ravel(order="F")
    return f


# The names a string coefficient may use (qutip/core/coefficient.py:367-395).
STR_ENV = {
    "sin": np.sin, "cos": np.cos, "tan": np.tan, "asin": np.arcsin, "acos": np.arccos,
    "atan": np.arctan, "pi": np.pi, "sinh": np.sinh, "cosh": np.cosh, "tanh": np.tanh,
    "asinh": np.arcsinh, "acosh": np.arccosh, "atanh": np.arctanh, "exp": np.exp,
    "log": np.log, "log10": np.log10, "sqrt": np.sqrt, "real": np.real, "imag": np.imag,
    "conj": np.conj, "abs": np.abs, "norm": lambda x: np.abs(x) ** 2, "arg": np.angle,
    "np": np,
}


def str_coefficient(base, args=None):
    """A string coefficient as StrFunctionCoefficient evaluates it
    (qutip/core/cy/coefficient.pyx:292-370): the expression in ``t`` and the
    arguments, run by Python in the environment above; returns f(t) ->
    complex."""
    args = dict(args or {})
    env = dict(STR_ENV)

    def f(t):
        return complex(eval(base, env, {"t": t, **args}))
    return f


def td_rhs(A0, parts):
    """dy/dt = (A0 + sum_e c_e(t) A_e) y: the list-form QobjEvo's
    matmul_data (qutip/core/cy/qobjevo.pyx:1103-1116 over _EvoElement,
    _element.pyx:117-182, 388-446), every coefficient a function of t."""
    A0 = sp.csr_matrix(A0, dtype=CD)
    parts = [(sp.csr_matrix(A, dtype=CD), c) for A, c in parts]

    def f(t, y):
        out = A0 @ y
        for A, c in parts:
            out = out + c(t) * (A @ y)
        return out
    return f


def vec(rho):
    """Column stacking, superoperator.py:291-302."""
    return np.asarray(rho).ravel(order="F")


def unvec(v, d):
    return np.asarray(v).reshape((d, d), order="F")


def ket2dm(psi):
    return np.outer(psi, psi.conj())


# ---------------------------------------------------------------------------
# benchmark systems
# ---------------------------------------------------------------------------
def jaynes_cummings(N=10, wc=2 * np.pi, wq=2 * np.pi, g=0.05 * 2 * np.pi,
                    kappa=0.005, gamma=0.05):
    """Cavity (N) x qubit, d = 2N = 20; Liouvillian 400 x 400."""
    a = sp.kron(destroy(N), qeye(2), format="csr")
    sm = sp.kron(qeye(N), sigmam(), format="csr")
    sz = sp.kron(qeye(N), sigmaz(), format="csr")
    H = wc * (dag(a) @ a) + wq * 0.5 * sz + g * (a @ dag(sm) + dag(a) @ sm)
    c_ops = [np.sqrt(kappa) * a, np.sqrt(gamma) * sm]
    psi0 = np.kron(basis(N, 0), basis(2, 0))
    e_ops = [dag(a) @ a, sz]
    return dict(H=sp.csr_matrix(H), c_ops=c_ops, psi0=psi0, e_ops=e_ops,
                tlist=np.linspace(0, 25, 100))


def ising_fields(L, seed=0):
    return np.random.default_rng(seed).uniform(0.5, 1.5, L)


def ising(L=8, gamma=0.1, J=1.0, seed=0, tf=10.0, npts=200):
    """Dissipative transverse-field Ising chain, open boundary.

    L=8: d=256, Liouvillian 65536^2 with 1 245 183 non-zeros.
    L=10 (gamma=0.05, t in [0, 20], 100 points): the mcsolve system, d=1024.
    """
    h = ising_fields(L, seed)
    d = 2 ** L
    H = sp.csr_matrix((d, d), dtype=CD)
    for i in range(L - 1):
        H = H + J * (site_op(sigmaz(), i, L) @ site_op(sigmaz(), i + 1, L))
    for i in range(L):
        H = H + h[i] * site_op(sigmax(), i, L)
    c_ops = [np.sqrt(gamma) * site_op(sigmam(), i, L) for i in range(L)]
    psi0 = kron_all_vec([basis(2, 0)] * L)
    mz = sum(site_op(sigmaz(), i, L) for i in range(L)) / L
    zz = site_op(sigmaz(), 0, L) @ site_op(sigmaz(), 1, L)
    return dict(H=sp.csr_matrix(H), c_ops=c_ops, psi0=psi0,
                e_ops=[sp.csr_matrix(mz), sp.csr_matrix(zz)],
                tlist=np.linspace(0, tf, npts), h=h)


def kron_all_vec(vs):
    out = np.ones(1, dtype=CD)
    for v in vs:
        out = np.kron(out, v)
    return out


def kerr(N=256, delta=0.5, K=0.1, F=2.0, kappa=1.0):
    """Driven-dissipative Kerr oscillator, dense rho N x N."""
    a = destroy(N)
    ad = dag(a)
    H = delta * (ad @ a) + 0.5 * K * (ad @ ad @ a @ a) + F * (a + ad)
    c_ops = [np.sqrt(kappa) * a]
    psi0 = basis(N, 0)
    parity = sp.diags((-1.0) ** np.arange(N), 0, format="csr", dtype=CD)
    return dict(H=sp.csr_matrix(H), c_ops=c_ops, psi0=psi0,
                e_ops=[sp.csr_matrix(ad @ a), parity],
                tlist=np.linspace(0, 30, 150))


def expect_dm(O, rho):
    """tr(O rho), qutip/core/data/expect.pyx:146-158."""
    return complex((O.multiply(np.asarray(rho).T)).sum())


def expect_ket(O, psi):
    """<psi|O|psi>, qutip/core/data/expect.pyx:133-144."""
    return complex(np.vdot(psi, O @ psi))
