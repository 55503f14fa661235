"""ctypes binding of libcntrx.so (the C ABI declared in include/cntrx.h)."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libcntrx.so")

CNTRX_OK, CNTRX_EINVAL, CNTRX_ENOMEM, CNTRX_EHIP, CNTRX_EUNSUPPORTED = 0, -1, -2, -3, -4
CNTRX_MAT_FOURIER, CNTRX_MAT_CG, CNTRX_MAT_FIXPOINT = 0, 1, 2


class CntrxHerm(ctypes.Structure):
    _fields_ = [("nt", ctypes.c_int), ("ntau", ctypes.c_int), ("size1", ctypes.c_int),
                ("sig", ctypes.c_int), ("ret", ctypes.c_void_p), ("les", ctypes.c_void_p),
                ("tv", ctypes.c_void_p), ("mat", ctypes.c_void_p)]


class CntrxFunc(ctypes.Structure):
    _fields_ = [("nt", ctypes.c_int), ("size1", ctypes.c_int), ("data", ctypes.c_void_p)]


class CntrxRequest(ctypes.Structure):
    _fields_ = [("size1", ctypes.c_int), ("nt", ctypes.c_int), ("ntau", ctypes.c_int),
                ("sig", ctypes.c_int), ("SolveOrder", ctypes.c_int), ("mat_method", ctypes.c_int),
                ("beta", ctypes.c_double), ("h", ctypes.c_double), ("mu", ctypes.c_double),
                ("V", ctypes.c_double), ("H", ctypes.c_void_p), ("Hbath", ctypes.c_void_p)]


class CntrxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("cntrx error %d: %s" % (code, msg))
        self.code = code


# exported symbol -> argument types (all return int unless listed in _DOUBLE)
_H = ctypes.POINTER(CntrxHerm)
_F = ctypes.POINTER(CntrxFunc)
_i, _d, _vp = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
SIGNATURES = {
    "cntrx_init": [_i],
    "cntrx_last_error": [],
    "cntrx_wait_timeouts": [_vp],
    "cntrx_integrator_tables": [_i, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "cntrx_dyson_mat": [_H, _d, _F, _H, _d, _i, _i, _i, _vp],
    "cntrx_dyson_start": [_H, _d, _F, _H, _d, _d, _i, _vp],
    "cntrx_dyson_timestep": [_i, _H, _d, _F, _H, _d, _d, _i, _vp],
    "cntrx_dyson_timestep_noalloc": [_i, _H, _d, _F, _H, _d, _d, _i, _vp],
    "cntrx_autotune": [_i],
    "cntrx_autotune_decisions": [_i, _vp, _vp],
    "cntrx_dyson_timestep_batched": [_i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp],
    "cntrx_vie2_mat": [_H, _H, _H, _H, _d, _i, _i, _vp],
    "cntrx_vie2_start": [_H, _H, _H, _H, _d, _d, _i, _vp],
    "cntrx_vie2_timestep": [_i, _H, _H, _H, _H, _d, _d, _i, _i, _vp],
    "cntrx_convolution_timestep": [_i, _H, _H, _H, _H, _H, _d, _d, _i, _vp],
    "cntrx_incr_convolution": [_i, _vp, _d, _d, _H, _H, _H, _vp, _vp, _H, _H, _d, _d, _i, _vp],
    "cntrx_bubble1": [_i, _H, _i, _i, _H, _H, _i, _i, _H, _H, _i, _i, _vp],
    "cntrx_bubble2": [_i, _H, _i, _i, _H, _H, _i, _i, _H, _H, _i, _i, _vp],
    "cntrx_extrapolate_timestep": [_i, _H, _i, _vp],
    "cntrx_distance_norm2": [_i, _H, _H, _vp, _vp],
    "cntrx_green_from_H": [_H, _d, _vp, _d, _d, _vp],
    "cntrx_green_from_H_td": [_H, _d, _F, _d, _d, _i, _i, _i, _i, _vp],
    "cntrx_force_matsubara_hermitian": [_H, _vp],
    "cntrx_timestep_lincomb": [_i, _H, _i, _vp, _vp, _vp],
    "cntrx_timestep_scale_by_function": [_i, _H, _F, _d, _i, _vp],
    "cntrx_copy": [_vp, _vp, ctypes.c_size_t, _i, _vp],
    "cntrx_device_alloc": [ctypes.c_size_t, _vp],
    "cntrx_device_free": [_vp],
    "cntrx_herm_print_to_file": [_H, ctypes.c_char_p, _i, _vp],
    "cntrx_herm_file_header": [ctypes.c_char_p, _vp, _vp, _vp, _vp],
    "cntrx_herm_read_from_file": [_H, ctypes.c_char_p, _vp, _vp],
    "cntrx_herm_read_from_file_nt1": [_H, _i, ctypes.c_char_p, _vp],
    "cntrx_function_print_to_file": [_F, ctypes.c_char_p, _i, _vp],
    "cntrx_function_file_header": [ctypes.c_char_p, _vp, _vp],
    "cntrx_function_read_from_file": [_F, ctypes.c_char_p, _vp],
    "cntrx_function_read_from_file_nt1": [_F, _i, ctypes.c_char_p, _vp],
    "cntrx_lattice_create": [_i, _vp, _d, _d, _d, _i, _i, _d, _i, _vp],
    "cntrx_lattice_destroy": [_vp],
    "cntrx_lattice_matrix": [_vp, _i, _H],
    "cntrx_lattice_equilibrium": [_vp, _vp],
    "cntrx_lattice_bootstrap": [_vp, _d, _i, _vp, _vp],
    "cntrx_lattice_step": [_vp, _i, _i, _d, _i, _vp, _vp, _vp],
    "cntrx_lattice_run": [_vp, _i, _i, _d, _i, _vp, _vp, _vp],
    "cntrx_lattice_profile": [_vp, _i],
    "cntrx_lattice_kernel_stats": [_vp, _vp, _vp, _vp],
    "cntrx_lattice_kernel_samples": [_vp, ctypes.c_long, _vp, _vp, _vp, _vp],
    "cntrx_lattice_fused": [_vp, _vp],
    "cntrx_request_footprint": [_vp],
    "cntrx_server_create": [ctypes.c_size_t, _vp],
    "cntrx_server_destroy": [_vp],
    "cntrx_server_submit": [_vp, _vp, _vp],
    "cntrx_server_step": [_vp, _vp, _vp],
    "cntrx_server_poll": [_vp, ctypes.c_long, _vp, _vp, _vp],
    "cntrx_server_drain": [_vp],
    "cntrx_server_result": [_vp, ctypes.c_long, _H],
    "cntrx_server_release": [_vp, ctypes.c_long],
    "cntrx_server_stats": [_vp, _vp],
}
_RESTYPE = {"cntrx_last_error": ctypes.c_char_p, "cntrx_request_footprint": ctypes.c_size_t}

_lib = None


def load():
    """Load the in-tree HIP library; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError("libcntrx.so not built: run `python -m nessi_amd.build` (%s)" % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, ctypes.c_int)
        _lib = lib
    return _lib


def check(code):
    if code != CNTRX_OK:
        raise CntrxError(code, load().cntrx_last_error().decode())
    return code
