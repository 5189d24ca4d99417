"""ctypes binding of the C ABI declared in include/opflow_hip.h.

The library is loaded from opflow_amd/lib/libopflow_hip.so (built in-tree by
__graft_entry__.build()). There is no fallback: if the library is missing the
import of anything that needs it raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libopflow_hip.so")

OPF_OK = 0
OPF_ERR_ARG = -1
OPF_ERR_HIP = -2
OPF_ERR_UNSUPPORTED = -3
OPF_ERR_RANGE = -4
OPF_ERR_NOMEM = -5
OPF_ERR_NOTREADY = -6

OPF_F64, OPF_F32 = 0, 1
# device scratch of opf_ftcs_noalloc (OPF_FTCS_NOALLOC_MAX_BYTES in the header)
OPF_FTCS_NOALLOC_MAX_BYTES = 64 << 20


class OpfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"opflow_hip error {code}: {msg}")
        self.code = code


class FieldDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("dtype", C.c_int),
        ("loc", C.c_int * 3),
        ("bc_type", (C.c_int * 2) * 3),
        ("bc_value", (C.c_double * 2) * 3),
        ("ext", (C.c_int * 2) * 3),
        ("padding", C.c_int),
        ("mesh_start", C.c_int * 3), ("mesh_end", C.c_int * 3),
        ("local_start", C.c_int * 3), ("local_end", C.c_int * 3),
        ("assign_start", C.c_int * 3), ("assign_end", C.c_int * 3),
        ("access_start", C.c_int * 3), ("access_end", C.c_int * 3),
        ("logical_start", C.c_int * 3), ("logical_end", C.c_int * 3),
        ("offset", C.c_int * 3),
        ("extent", C.c_int * 3),
    ]


class ExprNode(C.Structure):
    _fields_ = [("kind", C.c_int), ("a", C.c_int), ("b", C.c_int), ("dim", C.c_int), ("value", C.c_double)]


class SolveParams(C.Structure):
    _fields_ = [
        ("solver", C.c_int), ("precond", C.c_int), ("tol", C.c_double), ("max_iter", C.c_int),
        ("static_mat", C.c_int), ("pin_value", C.c_int), ("precision", C.c_int),
        ("pre_smooth", C.c_int), ("post_smooth", C.c_int), ("verbose", C.c_int),
    ]


class EqnDesc(C.Structure):
    _fields_ = [("nodes", C.POINTER(ExprNode)), ("nnodes", C.c_int), ("lhs_root", C.c_int), ("rhs_root", C.c_int),
                ("fields", C.POINTER(C.c_void_p)), ("nfields", C.c_int)]


class SolveState(C.Structure):
    _fields_ = [("niter", C.c_int), ("relerr", C.c_double), ("abserr", C.c_double),
                ("nrefine", C.c_int), ("converged", C.c_int)]


class PoissonDesc(C.Structure):
    _fields_ = [("dim", C.c_int), ("n", C.c_int * 3), ("h", C.c_double * 3),
                ("bc", (C.c_int * 2) * 3), ("sigma", C.c_double)]


class SrvConfig(C.Structure):
    _fields_ = [("mem_budget", C.c_size_t), ("max_inflight", C.c_int), ("precision", C.c_int),
                ("batched", C.c_int), ("tol", C.c_double), ("max_iter", C.c_int)]


class SrvRequest(C.Structure):
    _fields_ = [("desc", PoissonDesc), ("seed", C.c_uint64), ("nbumps", C.c_int), ("noise", C.c_double),
                ("rhs", C.POINTER(C.c_double)), ("out_x", C.POINTER(C.c_double)), ("t_submit", C.c_double)]


class SrvResult(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_int), ("state", SolveState),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
                ("checksum", C.c_double)]


_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)

# name -> (restype, argtypes)
SIGNATURES = {
    "opf_ctx_create": (_I, [_I, C.POINTER(_P)]),
    "opf_ctx_destroy": (_I, [_P]),
    "opf_ctx_sync": (_I, [_P]),
    "opf_ctx_stream": (_P, [_P]),
    "opf_ctx_set_jit": (_I, [_P, _I]),
    "opf_last_error": (C.c_char_p, []),
    "opf_version": (C.c_char_p, []),
    "opf_mesh_uniform_arrays": (_I, [_I, _D, _D, _I, _DP, _DP]),
    "opf_mesh_create": (_I, [_P, _I, _IP, _IP, _I, _DP * 3, _DP * 3, C.POINTER(_P)]),
    "opf_mesh_destroy": (_I, [_P]),
    "opf_field_desc_finalize": (_I, [C.POINTER(FieldDesc), _IP, _IP]),
    "opf_field_create": (_I, [_P, C.POINTER(FieldDesc), C.POINTER(_P)]),
    "opf_field_destroy": (_I, [_P]),
    "opf_field_get_desc": (C.POINTER(FieldDesc), [_P]),
    "opf_field_nbytes": (C.c_size_t, [_P]),
    "opf_field_device_ptr": (_P, [_P]),
    "opf_field_upload": (_I, [_P, _P, C.c_size_t]),
    "opf_field_download": (_I, [_P, _P, C.c_size_t]),
    "opf_field_update_padding": (_I, [_P]),
    "opf_field_assign_const": (_I, [_P, _I, _D]),
    "opf_field_copy": (_I, [_P, _P]),
    "opf_field_swap": (_I, [_P, _P]),
    "opf_sweep": (_I, [C.POINTER(ExprNode), _I, _I, C.POINTER(_P), _I, _P, _I]),
    "opf_ftcs": (_I, [_P, _D, _I]),
    "opf_ftcs_blocked": (_I, [_P, _D, _I, _I]),
    "opf_ftcs_plan": (_I, [_P, _I, _I, _IP, _IP]),
    "opf_ftcs_noalloc": (_I, [_P, _D, _I]),
    "opf_cache_stats": (_I, [_P, C.POINTER(C.c_longlong)]),
    "opf_reduce": (_I, [_P, _I, _IP, _IP, _DP]),
    "opf_buf_create": (_I, [_P, C.c_size_t, C.POINTER(_P)]),
    "opf_buf_destroy": (_I, [_P, _P]),
    "opf_buf_upload": (_I, [_P, _P, _P, C.c_size_t]),
    "opf_buf_download": (_I, [_P, _P, _P, C.c_size_t]),
    "opf_buf_zero": (_I, [_P, _P, C.c_size_t]),
    "opf_gen_uniform": (_I, [C.c_uint64, C.c_size_t, _D, _D, _DP]),
    "opf_gen_bumps": (_I, [_P, C.POINTER(FieldDesc), C.c_uint64, _I, _D, _DP]),
    "opf_eqn_create": (_I, [C.POINTER(ExprNode), _I, _I, _I, C.POINTER(_P), _I, _P, C.POINTER(SolveParams),
                            C.POINTER(_P)]),
    "opf_eqn_solve": (_I, [_P, C.POINTER(SolveState)]),
    "opf_eqn_csr": (_I, [_P, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), _P, _P, _P, _P]),
    "opf_eqn_destroy": (_I, [_P]),
    "opf_ctx_set_stream": (_I, [_P, _P]),
    "opf_field_write_cart": (_I, [_P, C.c_char_p, C.c_char_p, _I, C.c_double]),
    "opf_field_read_cart": (_I, [_P, C.c_char_p, C.c_char_p, _I, C.POINTER(C.c_double)]),
    "opf_eqns_solve": (_I, [_I, C.POINTER(EqnDesc), C.POINTER(_P), C.POINTER(SolveParams), C.POINTER(SolveState)]),
    "opf_eqns_csr": (_I, [_I, C.POINTER(EqnDesc), C.POINTER(_P), _P, C.POINTER(C.c_longlong),
                          C.POINTER(C.c_longlong), _P, _P, _P, _P]),
    "opf_gmg_create": (_I, [_P, C.POINTER(PoissonDesc), _I, C.POINTER(_P)]),
    "opf_gmg_solve": (_I, [_P, _P, _P, C.POINTER(SolveParams), C.POINTER(SolveState)]),
    "opf_gmg_apply": (_I, [_P, _P, _P]),
    "opf_gmg_destroy": (_I, [_P]),
    "opf_gmg_footprint": (C.c_size_t, [C.POINTER(PoissonDesc), _I]),
    "opf_srv_create": (_I, [_P, C.POINTER(SrvConfig), C.POINTER(_P)]),
    "opf_srv_submit": (_I, [_P, C.POINTER(SrvRequest), C.POINTER(C.c_uint64)]),
    "opf_srv_step": (_I, [_P, _IP]),
    "opf_srv_poll": (_I, [_P, C.POINTER(SrvResult), _I, _IP]),
    "opf_gen_bumps_grid": (_I, [C.POINTER(PoissonDesc), C.c_uint64, _I, _D, _DP]),
    "opf_gen_bumps_grid_dev": (_I, [_P, C.POINTER(PoissonDesc), C.c_uint64, _I, _D, _P]),
    "opf_srv_pending": (_I, [_P, _IP, _IP]),
    "opf_srv_clock": (_D, [_P]),
    "opf_srv_destroy": (_I, [_P]),
}

_lib = None


def load(path=None):
    """Load libopflow_hip.so and bind every symbol of the C ABI."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"libopflow_hip.so not found at {p}: run __graft_entry__.build() first "
                           "(there is no CPU fallback)")
    lib = C.CDLL(p)
    missing = []
    for name, (res, args) in SIGNATURES.items():
        try:
            fn = getattr(lib, name)
        except AttributeError:
            missing.append(name)
            continue
        fn.restype = res
        fn.argtypes = args
    lib.missing_symbols = missing
    if path is None:
        _lib = lib
    return lib


def check(rc):
    if rc != OPF_OK:
        msg = load().opf_last_error()
        raise OpfError(rc, msg.decode() if msg else "")
    return rc
