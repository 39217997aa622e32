"""The C-ABI library: it loads and exports every symbol include/*.h
declares (no compute call: this runs without a GPU)."""
import ctypes
import pathlib
import re

import pytest

from qutip_amd import _lib

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(qa_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "qa_rk_run" in names and "qa_csr_matvec_z" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_symbols() if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_symbols()) <= set(_lib.SIGNATURES)


def test_load_sets_signatures_without_gpu():
    L = _lib.load()
    assert L.qa_version() == 1
    assert L.qa_last_error() is not None


def test_missing_library_fails_loudly(tmp_path):
    import importlib
    mod = importlib.reload(_lib)
    try:
        with pytest.raises(mod.QutipAmdUnavailable):
            mod.load(tmp_path / "nope.so")
    finally:
        importlib.reload(_lib)


def _tableau_args():
    """vern7 as qa_rk_create* take it (host arrays kept alive by the caller)."""
    import numpy as np
    from qutip_amd.solver.tableaus import TABLEAUS
    t = TABLEAUS["vern7"]
    arrs = [np.ascontiguousarray(t[k], dtype=np.float64) for k in ("a", "b", "c", "e", "bi")]
    dp = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    args = (int(t["order"]), len(arrs[1]), len(arrs[2]), arrs[4].shape[1], *map(dp, arrs),
            1e-8, 1e-10, 1000, 0.0, 0.0, 0.0, 1)
    return arrs, args


def test_coded_entry_points_reject_bad_splits():
    """The coded operator's code split must be 6..10 value bits; the checks
    run before any device call."""
    L = _lib.load()
    vp = ctypes.c_void_p
    rc = L.qa_sell_coded_matvec_z(64, 64, 1, 1, vp(16), vp(16), vp(16), vp(16), 5, vp(16),
                                  vp(16), 1.0, 0.0, 0, None)
    assert rc == _lib.QA_ERR_ARG and b"vbits" in L.qa_last_error()
    keep, targs = _tableau_args()
    h = vp()
    rc = L.qa_rk_create_coded(ctypes.byref(h), 64, 1, 1, vp(16), vp(16), vp(16), vp(16), 11,
                              *targs)
    assert rc == _lib.QA_ERR_ARG and b"vbits" in L.qa_last_error() and not h.value


def test_banded_entry_point_rejects_bad_diagonals():
    """qa_rk_create_dia: a diagonal of a missing operator, an offset of N, or
    more than 96 products of diagonals are argument errors, raised before
    any device call."""
    import numpy as np
    L = _lib.load()
    keep, targs = _tableau_args()
    N = 128
    ip = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    dp = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

    def create(ops, offs, nc):
        ops = np.asarray(ops, np.int32)
        offs = np.asarray(offs, np.int32)
        vals = np.ones((len(ops), N), np.complex128)
        h = ctypes.c_void_p()
        rc = L.qa_rk_create_dia(ctypes.byref(h), N, nc, len(ops), ip(ops), ip(offs),
                                dp(vals.view(np.float64)), *targs)
        return rc, h

    rc, h = create([0, 2], [0, 1], 1)  # operator 2 with one collapse operator
    assert rc == _lib.QA_ERR_ARG and not h.value
    rc, h = create([0], [N], 0)  # offset N lies outside the matrix
    assert rc == _lib.QA_ERR_ARG and not h.value
    rc, h = create([0] * 50, list(range(-25, 25)), 0)  # 100 products of H_nh's diagonals
    assert rc == _lib.QA_ERR_ARG and b"products of diagonals" in L.qa_last_error()


def _plan(d, h_width=0, nfun=0, ntimes=0, fp32=0, small=1):
    out = (ctypes.c_int64 * 8)()
    _lib.host_call("qa_mc_lds_plan", d, h_width, nfun, ntimes, fp32, small, out)
    return dict(zip(["threads", "rows", "reg_stages", "lds", "static", "hl", "fun_lds", "cap"],
                    list(out)))


@pytest.mark.parametrize("d,w,fp32", [(512, 8, 0), (512, 8, 1), (256, 12, 0), (400, 8, 0),
                                      (1024, 12, 0), (1024, 12, 1), (64, 4, 0), (20, 4, 0)])
def test_lds_plan_counts_the_static_tableau(d, w, fp32):
    """Every launch shape keeps dynamic + static LDS inside the CU's
    160 KiB (half of it for the two-workgroup shape), with the whole-solve
    room for functionals and times and the operator in LDS; the case that
    overflowed once is d = 512 at ELL width 8 with functionals."""
    budget = 160 * 1024
    for nfun, ntimes in ((0, 0), (3, 100), (40, 2000)):
        p = _plan(d, w, nfun, ntimes, fp32, 1)
        share = budget // 2 if p["reg_stages"] == 0 else budget
        assert p["static"] > 0
        assert p["lds"] + p["static"] <= share, p
        assert p["lds"] <= p["cap"]
        if p["fun_lds"]:
            assert p["fun_lds"] >= 256 and p["fun_lds"] % 16 == 0
        # the room for functionals never lowers the workgroups a CU holds
        base = _plan(d, w, 0, 0, fp32, 1)
        assert budget // (p["lds"] + p["static"]) == budget // (base["lds"] + base["static"])
    t = _plan(d, w, 3, 100, fp32, 0)  # the trajectory kernel reserves nothing extra
    assert t["fun_lds"] == 0 and t["hl"] == 0


def test_lds_plan_rejects_oversize():
    with pytest.raises(ValueError):
        _plan(4096)


def test_functionals_check():
    i32 = lambda *v: (ctypes.c_int * max(len(v), 1))(*v)
    ok = _lib.host_call("qa_functionals_check", 4, 2, i32(0, 1), i32(0, 2, 3), i32(0, 0, 3),
                        i32(1, 3, 2))
    assert ok == 0
    with pytest.raises(ValueError, match="column"):
        _lib.host_call("qa_functionals_check", 4, 1, i32(0), i32(0, 1), i32(0), i32(4))
    with pytest.raises(ValueError, match="row"):
        _lib.host_call("qa_functionals_check", 4, 1, i32(1), i32(0, 1), i32(-1), i32(0))
    with pytest.raises(ValueError, match="fptr"):
        _lib.host_call("qa_functionals_check", 4, 1, i32(0), i32(1, 1), i32(0), i32(0))


def _coef_check(progs, ntimes=0):
    """qa_coef_eval with no times: only the host-side program checks run."""
    import numpy as np
    from qutip_amd import coefficient as qc
    start, code, consts = qc.DeviceProgram.concat(progs).arrays()
    L = _lib.load()
    return L.qa_coef_eval(len(progs), start.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                          len(code), code.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                          len(consts),
                          consts.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                          ntimes, None, None, None)


def test_coefficient_programs_pass_the_abi_checks():
    """Every compiled test expression and interpolation table is accepted
    by the library's program checks; malformed programs are rejected
    before anything reaches the device."""
    import numpy as np
    from qutip_amd import coefficient as qc
    from tests.test_coefficient import EXPRESSIONS
    progs = [qc.coefficient(e, args=a).device_program() for e, a in EXPRESSIONS]
    tl = np.linspace(0, 1, 6)
    progs += [qc.coefficient(tl ** 2, tlist=tl, order=k).device_program() for k in (0, 1, 3)]
    for g0 in range(0, len(progs), 8):
        assert _coef_check(progs[g0:g0 + 8]) == 0, _lib.load().qa_last_error()
    assert _coef_check(progs[:9]) == _lib.QA_ERR_ARG  # more than 8 programs
    bad = [qc.DeviceProgram([qc.CO_ADD], [])]  # pops an empty stack
    assert _coef_check(bad) == _lib.QA_ERR_ARG
    bad = [qc.DeviceProgram([qc.CO_CONST | (3 << 8)], [1.0])]  # constant out of range
    assert _coef_check(bad) == _lib.QA_ERR_ARG
    bad = [qc.DeviceProgram([qc.CO_T, qc.CO_T], [])]  # leaves two values
    assert _coef_check(bad) == _lib.QA_ERR_ARG
    bad = [qc.DeviceProgram([qc.CO_INTERP], [complex(50, 2), 0j])]  # table past the end
    assert _coef_check(bad) == _lib.QA_ERR_ARG


def test_hermitian_switch_rejects_null_solver():
    """qa_rk_set_hermitian / qa_rk_get_hermitian check their arguments
    before touching a solver (no device call)."""
    L = _lib.load()
    assert L.qa_rk_set_hermitian(None, 1) == _lib.QA_ERR_ARG
    assert b"qa_rk_set_hermitian" in L.qa_last_error()
    on = ctypes.c_int(7)
    assert L.qa_rk_get_hermitian(None, ctypes.byref(on)) == _lib.QA_ERR_ARG
    assert on.value == 7


def test_sell_matmat_rejects_bad_shapes():
    """qa_sell_matmat_z checks lanes, slice count and leading dimensions
    before any device call."""
    L = _lib.load()
    vp = ctypes.c_void_p
    rc = L.qa_sell_matmat_z(64, 64, 3, 1, vp(16), vp(16), vp(16), 0, 2, vp(16), 64, vp(16), 64,
                            1.0, 0.0, 0, None)
    assert rc == _lib.QA_ERR_ARG and b"lanes" in L.qa_last_error()
    rc = L.qa_sell_matmat_z(65, 64, 1, 1, vp(16), vp(16), vp(16), 0, 2, vp(16), 64, vp(16), 65,
                            1.0, 0.0, 0, None)
    assert rc == _lib.QA_ERR_ARG and b"nslices" in L.qa_last_error()
    rc = L.qa_sell_matmat_z(64, 64, 1, 1, vp(16), vp(16), vp(16), 0, 2, vp(16), 63, vp(16), 64,
                            1.0, 0.0, 0, None)
    assert rc == _lib.QA_ERR_ARG and b"leading" in L.qa_last_error()


def test_dense_shape_rejects_null_solver():
    L = _lib.load()
    assert L.qa_rk_set_dense_shape(None, 0, 1, 1) == _lib.QA_ERR_ARG
    t, n, h = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    assert L.qa_rk_get_dense_shape(None, ctypes.byref(t), ctypes.byref(n),
                                   ctypes.byref(h)) == _lib.QA_ERR_ARG


def test_dense_shape_candidates():
    """The autotuner's candidate shapes are valid for N (no device)."""
    from qutip_amd import autotune
    c = autotune.dense_candidates(512, True, True, (1, 8, 8))
    assert c[0] == (1, 8, 8)
    assert (0, 8, 1) in c and (0, 8, 8) in c and len(set(c)) == len(c)
    c = autotune.dense_candidates(96, True, False, (0, 2, 1))
    assert all(t == 0 for t, _, _ in c)                 # 96 % 64 != 0: no LDS tiles
    assert {ns for _, ns, _ in c} == {1, 2}             # 96 % 32 == 0, 96 % 64 != 0
    c = autotune.dense_candidates(256, False, False, (0, 8, 8))
    assert {ns for t, ns, _ in c if t == 0} == {1, 2, 4, 8}   # bf16: 32-wide steps
