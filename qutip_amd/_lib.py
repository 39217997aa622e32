"""ctypes binding of ``libqutip_amd.so`` (declarations: include/qutip_amd.h).

The product has no CPU fallback: when the library is missing, or no GPU is
visible, every device operation raises ``QutipAmdUnavailable``.
"""
import ctypes as C
import os
import pathlib

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "libqutip_amd.so"

QA_OK = 0
QA_ERR_HIP = -1
QA_ERR_ARG = -2
QA_ERR_NOMEM = -3
QA_ERR_STATE = -4
QA_ERR_LIMIT = -5

vp = C.c_void_p
i32 = C.c_int
i64 = C.c_int64
f64 = C.c_double
P_i32 = C.POINTER(C.c_int)
P_f64 = C.POINTER(C.c_double)
P_i64 = C.POINTER(C.c_int64)

# name -> argtypes (every function returns int unless listed in _RESTYPES)
SIGNATURES = {
    "qa_last_error": [],
    "qa_version": [],
    "qa_device_count": [P_i32],
    "qa_set_device": [i32],
    "qa_device_info": [i32, C.c_char_p, i32, P_i32, C.POINTER(C.c_size_t)],
    "qa_mem_info": [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)],
    "qa_malloc": [C.POINTER(vp), C.c_size_t],
    "qa_free": [vp],
    "qa_host_malloc": [C.POINTER(vp), C.c_size_t],
    "qa_host_free": [vp],
    "qa_memcpy_h2d": [vp, vp, C.c_size_t, vp],
    "qa_memcpy_d2h": [vp, vp, C.c_size_t, vp],
    "qa_memcpy_d2d": [vp, vp, C.c_size_t, vp],
    "qa_fill_zero": [vp, C.c_size_t, vp],
    "qa_stream_create": [C.POINTER(vp)],
    "qa_stream_destroy": [vp],
    "qa_stream_sync": [vp],
    "qa_device_sync": [],
    "qa_event_create": [C.POINTER(vp)],
    "qa_event_destroy": [vp],
    "qa_event_record": [vp, vp],
    "qa_event_query": [vp, P_i32],
    "qa_csr_matvec_z": [i32, i32, vp, vp, vp, i32, i32, vp, vp, f64, f64, i32, vp],
    "qa_csr_matmat_z": [i32, i32, vp, vp, vp, i32, i32, vp, i64, vp, i64, f64, f64, i32, vp],
    "qa_sell_matmat_z": [i32, i32, i32, i32, vp, vp, vp, i32, i32, vp, i64, vp, i64, f64, f64, i32,
                         vp],
    "qa_sell_matvec_z": [i32, i32, i32, vp, vp, vp, i32, vp, vp, f64, f64, i32, vp],
    "qa_sell_coded_matvec_z": [i32, i32, i32, i32, vp, vp, vp, vp, i32, vp, vp, f64, f64, i32,
                               vp],
    "qa_lincomb_z": [i64, vp, i32, C.POINTER(vp), P_f64, vp, vp],
    "qa_axpby_z": [i64, f64, f64, vp, f64, f64, vp, vp, vp],
    "qa_wrmn_error_z": [i64, vp, vp, f64, f64, P_f64, vp],
    "qa_wrmn_error_z_noalloc": [i64, vp, vp, f64, f64, P_f64, vp],
    "qa_norm2_z": [i64, vp, P_f64, vp],
    "qa_inner_z": [i64, vp, vp, P_f64, vp],
    "qa_trace_z": [i32, vp, P_f64, vp],
    "qa_expect_csr_ket_z": [i32, vp, vp, vp, i32, vp, P_f64, vp],
    "qa_expect_csr_dm_z": [i32, vp, vp, vp, i32, vp, P_f64, vp],
    "qa_rk_create": [C.POINTER(vp), i32, i32, i32, vp, vp, vp, i32, i32, i32, i32, i32,
                     P_f64, P_f64, P_f64, P_f64, P_f64, f64, f64, i32, f64, f64, f64, i32],
    "qa_rk_create_coded": [C.POINTER(vp), i32, i32, i32, vp, vp, vp, vp, i32, i32, i32, i32, i32,
                           P_f64, P_f64, P_f64, P_f64, P_f64, f64, f64, i32, f64, f64, f64, i32],
    "qa_rk_create_dia": [C.POINTER(vp), i32, i32, i32, P_i32, P_i32, P_f64, i32, i32, i32, i32,
                         P_f64, P_f64, P_f64, P_f64, P_f64, f64, f64, i32, f64, f64, f64, i32],
    "qa_rk_create_dense": [C.POINTER(vp), i32, i32, P_f64, P_f64, i32, i32, i32, i32, i32,
                           P_f64, P_f64, P_f64, P_f64, P_f64, f64, f64, i32, f64, f64, f64, i32],
    "qa_rk_set_hermitian": [vp, i32],
    "qa_rk_get_hermitian": [vp, P_i32],
    "qa_rk_set_dense_shape": [vp, i32, i32, i32],
    "qa_rk_get_dense_shape": [vp, P_i32, P_i32, P_i32],
    "qa_rk_apply": [vp, vp, vp],
    "qa_rk_apply_at": [vp, f64, vp, vp],
    "qa_rk_set_coefficients": [vp, i32, P_i32, P_i32, i32, P_i32, i32, P_f64],
    "qa_coef_eval": [i32, P_i32, i32, P_i32, i32, P_f64, i32, vp, vp, vp],
    "qa_rk_destroy": [vp],
    "qa_rk_stream": [vp],
    "qa_rk_set_functionals": [vp, i32, P_i32, P_i32, P_i32, P_i32, P_f64],
    "qa_rk_set_state": [vp, f64, vp],
    "qa_rk_run": [vp, i32, P_f64, i32, P_f64, vp, P_i32, P_f64],
    "qa_rk_begin": [vp, i32, P_f64, i32, vp],
    "qa_rk_advance": [vp, P_i32],
    "qa_rk_finish": [vp, P_f64, P_i32, P_f64],
    "qa_rk_get_state": [vp, P_f64, vp],
    "qa_rk_stats": [vp, P_i64],
    "qa_rk_iterate": [vp, i32],
    "qa_rk_profile_stages": [vp, i32, P_f64],
    "qa_rk_profile_sequence": [vp, i32, P_f64, P_f64],
    "qa_mc_tableau_size": [],
    "qa_mc_lds_plan": [i32, i32, i32, i32, i32, i32, P_i64],
    "qa_functionals_check": [i32, i32, P_i32, P_i32, P_i32, P_i32],
    "qa_mc_pack_tableau": [vp, i32, i32, i32, i32, P_f64, P_f64, P_f64, P_f64, P_f64, i32],
    "qa_rk_small_run": [i32, i32, f64, i32, vp, vp, i32, vp, vp, C.c_longlong, i32, vp, vp, vp,
                        vp, vp, vp, f64, f64, i32, f64, f64, f64, i32, vp, vp, vp, i32, vp],
    "qa_sde_run": [i32, i32, i32, i32, i32, i32, i32, vp, f64, C.c_longlong, i32, vp, vp, i32, vp,
                   vp, vp, vp, i32, vp, vp, vp, vp, vp, i32, vp],
    "qa_mc_run": [i32, i32, i32, vp, vp, i32, vp, vp, i32, vp, vp, vp, i32, vp, vp, vp, vp, vp,
                  i32, f64, f64, i32, f64, f64, f64, i32, f64, f64, f64, f64, i32, vp, vp, vp,
                  i32, vp, i32, vp],
}
_RESTYPES = {"qa_last_error": C.c_char_p, "qa_rk_stream": vp}


class QutipAmdUnavailable(RuntimeError):
    """The gfx950 library or a GPU is not available; there is no fallback."""


class QutipAmdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None
_gpu_checked = False


def load(path=None):
    """Load the shared library (does not touch the GPU)."""
    global _lib
    if _lib is not None:
        return _lib
    p = pathlib.Path(path or os.environ.get("QUTIP_AMD_LIB", LIB_PATH))
    if not p.exists():
        raise QutipAmdUnavailable(
            f"{p} is not built; run `python -m qutip_amd._build` (hipcc, gfx950)")
    lib = C.CDLL(str(p))
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, C.c_int)
    _lib = lib
    return lib


def exported_symbols():
    return list(SIGNATURES)


def lib():
    """The library, after checking once that a GPU is visible."""
    global _gpu_checked
    L = load()
    if not _gpu_checked:
        n = C.c_int(0)
        rc = L.qa_device_count(C.byref(n))
        if rc != QA_OK or n.value < 1:
            raise QutipAmdUnavailable(
                "no HIP device visible: qutip_amd runs only on the GPU "
                f"(qa_device_count rc={rc}, n={n.value})")
        _gpu_checked = True
    return L


def check(rc, what=""):
    if rc == QA_OK:
        return
    msg = _lib.qa_last_error().decode(errors="replace") if _lib is not None else ""
    if rc == QA_ERR_NOMEM:
        raise MemoryError(f"{what}: {msg}")
    if rc == QA_ERR_ARG:
        raise ValueError(f"{what}: {msg}")
    raise QutipAmdError(rc, f"{what}: {msg}")


def call(name, *args):
    rc = getattr(lib(), name)(*args)
    check(rc, name)
    return rc


def host_call(name, *args):
    """Call an entry point that does host arithmetic only (no GPU needed)."""
    rc = getattr(load(), name)(*args)
    check(rc, name)
    return rc
