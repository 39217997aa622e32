"""ctypes bindings of the C oracle (oracle/c/qt_oracle.c) and of the
reference's own spmv built into oracle/_ref (test infrastructure only)."""
import ctypes as C
import pathlib

import numpy as np
import scipy.sparse as sp

HERE = pathlib.Path(__file__).resolve().parent
LIB_ORACLE = HERE / "c" / "libqt_oracle.so"
LIB_REF = HERE / "_ref" / "libqtref_spmv.so"

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def _load(path):
    if not path.exists():
        raise FileNotFoundError(f"{path} is not built (run `make -C oracle`)")
    return C.CDLL(str(path))


_oracle = None
_ref = None


def oracle_lib():
    global _oracle
    if _oracle is None:
        lib = _load(LIB_ORACLE)
        lib.qto_csr_matvec.argtypes = [C.c_int, _ip, _ip, _dp, _dp, C.c_double,
                                       C.c_double, _dp]
        lib.qto_csr_matvec.restype = None
        lib.qto_rk_solve.restype = C.c_int
        _oracle = lib
    return _oracle


def ref_lib():
    global _ref
    if _ref is None:
        lib = _load(LIB_REF)
        lib.qtref_csr_matvec.argtypes = [_dp, _ip, _ip, _dp, C.c_double, C.c_double,
                                         _dp, C.c_int]
        lib.qtref_csr_matvec.restype = None
        _ref = lib
    return _ref


def ref_available():
    return LIB_REF.exists()


def _csr_parts(A):
    A = sp.csr_matrix(A, dtype=np.complex128)
    return (np.ascontiguousarray(A.indptr, dtype=np.int32),
            np.ascontiguousarray(A.indices, dtype=np.int32),
            np.ascontiguousarray(A.data).view(np.float64))


def csr_matvec(A, x, out=None, scale=1.0):
    """out += scale * A @ x with the C restatement."""
    rp, ci, dv = _csr_parts(A)
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros(A.shape[0], dtype=np.complex128) if out is None else out
    oracle_lib().qto_csr_matvec(A.shape[0], rp, ci, dv, x.view(np.float64),
                                float(np.real(scale)), float(np.imag(scale)),
                                out.view(np.float64))
    return out


def ref_csr_matvec(A, x, out=None, scale=1.0):
    """out += scale * A @ x with the reference's own C++ kernel."""
    rp, ci, dv = _csr_parts(A)
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros(A.shape[0], dtype=np.complex128) if out is None else out
    ref_lib().qtref_csr_matvec(dv, ci, rp, x.view(np.float64), float(np.real(scale)),
                               float(np.imag(scale)), out.view(np.float64), A.shape[0])
    return out


def functionals_from_ops(e_ops, d, with_trace=True):
    """Sparse weights w with tr(O rho) = sum_i w_i vec(rho)[i]:
    vec index c*d + r holds rho[r, c], and tr(O rho) = sum O[r, c] rho[c, r],
    so weight O[r, c] goes to index r*d + c."""
    ptr, idx, w = [0], [], []
    ops = list(e_ops)
    if with_trace:
        ops = ops + [sp.identity(d, dtype=np.complex128, format="csr")]
    for O in ops:
        O = sp.coo_matrix(O)
        idx.extend((O.row * d + O.col).tolist())
        w.extend(O.data.tolist())
        ptr.append(len(idx))
    return (np.asarray(ptr, dtype=np.int32), np.asarray(idx, dtype=np.int32),
            np.ascontiguousarray(np.asarray(w, dtype=np.complex128)).view(np.float64))


def rk_solve(A, y0, tlist, tableau, fun, rtol=1e-6, atol=1e-8, nsteps=1000,
             first_step=0.0, min_step=0.0, max_step=0.0, interpolate=True):
    """Run the C restatement of explicit_rk.pyx on dy/dt = A y.

    ``fun`` = (ptr, idx, w) functionals; returns (values[ntimes, nfun],
    y_final, stats dict)."""
    lib = oracle_lib()
    rp, ci, dv = _csr_parts(A)
    n = A.shape[0]
    y0 = np.ascontiguousarray(y0, dtype=np.complex128)
    tl = np.ascontiguousarray(tlist, dtype=np.float64)
    a = np.ascontiguousarray(tableau["a"], dtype=np.float64)
    b = np.ascontiguousarray(tableau["b"], dtype=np.float64)
    c = np.ascontiguousarray(tableau["c"], dtype=np.float64)
    e = np.ascontiguousarray(tableau["e"], dtype=np.float64)
    bi = np.ascontiguousarray(tableau["bi"], dtype=np.float64)
    fptr, fidx, fw = fun
    nfun = len(fptr) - 1
    out = np.zeros((len(tl), nfun), dtype=np.complex128)
    yf = np.zeros(n, dtype=np.complex128)
    stats = np.zeros(3, dtype=np.int64)
    P = lambda arr: arr.ctypes.data_as(C.c_void_p)
    st = lib.qto_rk_solve(
        C.c_int(n), P(rp), P(ci), P(dv), P(y0), C.c_int(len(tl)), P(tl),
        C.c_int(tableau["order"]), C.c_int(len(b)), C.c_int(len(c)), C.c_int(bi.shape[1]),
        P(a), P(b), P(c), P(e), P(bi),
        C.c_double(rtol), C.c_double(atol), C.c_int(nsteps), C.c_double(first_step),
        C.c_double(min_step), C.c_double(max_step), C.c_int(int(interpolate)),
        C.c_int(nfun), P(fptr), P(fidx), P(fw), P(out), P(yf), P(stats))
    return out, yf, {"n_rhs": int(stats[0]), "n_attempts": int(stats[1]), "status": int(st)}


def _csr_blocks(ops, d):
    """CSR operators back to back with absolute row pointers (one block of
    d + 1 per operator)."""
    rps, cis, dvs, off = [], [], [], 0
    for O in ops:
        A = sp.csr_matrix(O, dtype=np.complex128)
        A.sort_indices()
        rps.append(A.indptr.astype(np.int64) + off)
        cis.append(A.indices)
        dvs.append(A.data)
        off += A.nnz
    rp = np.ascontiguousarray(np.concatenate(rps), dtype=np.int32)
    ci = np.ascontiguousarray(np.concatenate(cis) if off else np.zeros(1), dtype=np.int32)
    dv = np.ascontiguousarray(np.concatenate(dvs) if off else np.zeros(1, np.complex128),
                              dtype=np.complex128).view(np.float64)
    return rp, ci, dv


class McTrajectories:
    """The C restatement of the mcsolve trajectory loop (qto_mc_traj) for
    one system: ``run(psi0, tlist, draws)`` gives <O>(t) of one trajectory
    from its uniform draws (numpy's stream of its generator)."""

    def __init__(self, H, c_ops, e_ops, tableau, rtol=1e-6, atol=1e-8, nsteps=1000,
                 norm_steps=25, norm_t_tol=1e-6, norm_tol=1e-4, norm_min_step=0.1,
                 mc_corr_eps=1e-10):
        H = sp.csr_matrix(H, dtype=np.complex128)
        cs = [sp.csr_matrix(c, dtype=np.complex128) for c in c_ops]
        ns = [(c.conj().T @ c).tocsr() for c in cs]
        Heff = -1j * H
        for n_ in ns:
            Heff = Heff - 0.5 * n_
        self.d = H.shape[0]
        self.h = _csr_parts(sp.csr_matrix(Heff))
        self.c = _csr_blocks(cs, self.d)
        self.n = _csr_blocks(ns, self.d)
        self.e = _csr_blocks(e_ops, self.d)
        self.nchan, self.neops = len(cs), len(e_ops)
        self.tab = [np.ascontiguousarray(tableau[k], dtype=np.float64)
                    for k in ("a", "b", "c", "e", "bi")]
        self.order = int(tableau["order"])
        self.opts = (rtol, atol, nsteps, norm_steps, norm_t_tol, norm_tol, norm_min_step,
                     mc_corr_eps)

    def run(self, psi0, tlist, draws):
        lib = oracle_lib()
        P = lambda arr: arr.ctypes.data_as(C.c_void_p)
        psi0 = np.ascontiguousarray(psi0, dtype=np.complex128)
        tl = np.ascontiguousarray(tlist, dtype=np.float64)
        draws = np.ascontiguousarray(draws, dtype=np.float64)
        out = np.zeros((len(tl), max(self.neops, 1)), dtype=np.complex128)
        stats = np.zeros(4, dtype=np.int64)
        a, b, c, e, bi = self.tab
        rtol, atol, nsteps, norm_steps, norm_t_tol, norm_tol, norm_min_step, eps = self.opts
        rc = lib.qto_mc_traj(
            C.c_int(self.d), P(self.h[0]), P(self.h[1]), P(self.h[2]), C.c_int(self.nchan),
            P(self.c[0]), P(self.c[1]), P(self.c[2]), P(self.n[0]), P(self.n[1]),
            P(self.n[2]), C.c_int(self.neops), P(self.e[0]), P(self.e[1]), P(self.e[2]),
            P(psi0), C.c_int(len(tl)), P(tl), C.c_int(self.order), C.c_int(len(b)),
            C.c_int(len(c)), C.c_int(bi.shape[1]), P(a), P(b), P(c), P(e), P(bi),
            C.c_double(rtol), C.c_double(atol), C.c_int(nsteps), C.c_int(norm_steps),
            C.c_double(norm_t_tol), C.c_double(norm_tol), C.c_double(norm_min_step),
            C.c_double(eps), P(draws), C.c_int(len(draws)), P(out), P(stats))
        return out[:, :self.neops], {"rc": int(rc), "n_rhs": int(stats[0]),
                                     "n_attempts": int(stats[1]), "jumps": int(stats[2]),
                                     "draws": int(stats[3])}


def diagonals(M, tol=0.0):
    """(offsets, values [nd, N]) of a square matrix with values[d][r] =
    M[r, r + o_d] (zero outside the matrix); the non-zero diagonals only."""
    M = np.asarray(M.toarray() if sp.issparse(M) else M, dtype=np.complex128)
    N = M.shape[0]
    offs, vals = [], []
    for o in range(-N + 1, N):
        dg = np.diagonal(M, o)
        if np.abs(dg).max() > tol:
            v = np.zeros(N, dtype=np.complex128)
            if o >= 0:
                v[:N - o] = dg
            else:
                v[-o:] = dg
            offs.append(o)
            vals.append(v)
    return np.asarray(offs, dtype=np.int32), np.asarray(vals).reshape(len(offs), N)


def rk_solve_dia(H, c_ops, rho0, tlist, tableau, fun, rtol=1e-8, atol=1e-10, nsteps=10 ** 6):
    """The C restatement of the RK on the banded matrix form (qto_rk_solve_dia):
    H_nh = H - i/2 sum c^dag c and the c_ops as diagonals.  Returns
    (values[ntimes, nfun], stats)."""
    lib = oracle_lib()
    H = np.asarray(H.toarray() if sp.issparse(H) else H, dtype=np.complex128)
    cs = [np.asarray(c.toarray() if sp.issparse(c) else c, dtype=np.complex128) for c in c_ops]
    Hnh = H - 0.5j * sum((c.conj().T @ c for c in cs), np.zeros_like(H))
    ho, hv = diagonals(Hnh)
    cd = [diagonals(c) for c in cs]
    c_nd = np.asarray([len(o) for o, _ in cd] or [0], dtype=np.int32)
    c_off = np.ascontiguousarray(np.concatenate([o for o, _ in cd]) if cd else np.zeros(1),
                                 dtype=np.int32)
    c_val = np.ascontiguousarray(np.concatenate([v for _, v in cd]) if cd else
                                 np.zeros((1, H.shape[0]), np.complex128))
    hv = np.ascontiguousarray(hv)
    N = H.shape[0]
    rho0 = np.ascontiguousarray(rho0, dtype=np.complex128)
    tl = np.ascontiguousarray(tlist, dtype=np.float64)
    a, b, c, e, bi = [np.ascontiguousarray(tableau[k], dtype=np.float64)
                      for k in ("a", "b", "c", "e", "bi")]
    fptr, fidx, fw = fun
    nfun = len(fptr) - 1
    out = np.zeros((len(tl), nfun), dtype=np.complex128)
    stats = np.zeros(3, dtype=np.int64)
    P = lambda arr: arr.ctypes.data_as(C.c_void_p)
    st = lib.qto_rk_solve_dia(
        C.c_int(N), C.c_int(len(ho)), P(ho), P(hv.view(np.float64)), C.c_int(len(cs)),
        P(c_nd), P(c_off), P(c_val.view(np.float64)), P(rho0), C.c_int(len(tl)), P(tl),
        C.c_int(tableau["order"]), C.c_int(len(b)), C.c_int(len(c)), C.c_int(bi.shape[1]),
        P(a), P(b), P(c), P(e), P(bi), C.c_double(rtol), C.c_double(atol), C.c_int(nsteps),
        C.c_int(nfun), P(fptr), P(fidx), P(fw), P(out), P(stats))
    return out, {"n_rhs": int(stats[0]), "n_attempts": int(stats[1]), "status": int(st)}
