"""ctypes binding of libcnpc_hip.so (the C ABI declared in include/cnpc.h).

The library is required: importing the GPU path without it raises, there is
no CPU fallback.  Device buffers are torch tensors on the current HIP device;
only their data pointers cross the boundary.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcnpc_hip.so")


class cnpc_poly(ctypes.Structure):
    """Mirror of ``cnpc_poly`` in include/cnpc.h."""

    _fields_ = [
        ("m", ctypes.c_int32),
        ("deg", ctypes.c_int32),
        ("part", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("zc", ctypes.c_double * 9),
        ("pre_re", ctypes.c_double),
        ("pre_im", ctypes.c_double),
        ("k0_re", ctypes.c_double),
        ("k0_im", ctypes.c_double),
        ("rx", ctypes.c_double),
        ("ry", ctypes.c_double),
        ("rz", ctypes.c_double),
    ]


_P = ctypes.c_void_p
_I = ctypes.c_int
_L = ctypes.c_int64
_D = ctypes.c_double
_F = ctypes.c_float

# name -> argtypes (all return int status)
SIGNATURES = {
    "cnpc_version": [],
    "cnpc_ctx_create": [ctypes.POINTER(_P), _I],
    "cnpc_ctx_destroy": [_P],
    "cnpc_ctx_set_stream": [_P, _P],
    "cnpc_ctx_reserve": [_P, _L],
    "cnpc_ctx_hold": [_P, _I],
    "cnpc_ctx_clear_plans": [_P],
    "cnpc_paint": [_P, _P, _P, _L, _I, _D, _I, _I, _P, _I, _P, _P],
    "cnpc_paint_batch": [_P, _P, _P, _P, _I, _L, _L, _I, _D, _I, _I, _P, _I, _P, _P],
    "cnpc_fft_r2c": [_P, _I, _I, _I, _P, _P],
    "cnpc_fft_c2r": [_P, _I, _I, _I, _P, _P],
    "cnpc_prepare_spectrum": [_P, _I, _D, _I, _P, _P, _I, _I, _I, _D, _D, _I, _P],
    "cnpc_bin_spectrum": [_P, _I, _D, _I, _P, _P, _P, _I, _I, _P, _D, _I, _D, _P],
    "cnpc_cross_spectrum": [_P, _I, _I, _P, _P, _D, _I, _D, _P],
    "cnpc_resample_spectrum": [_P, _I, _I, _I, _P, _I, _P],
    "cnpc_shell_filter": [_P, _I, _D, _I, _P, _P, _I, _I, _I, _P, _P],
    "cnpc_shell_synth": [_P, _I, _D, _I, _P, _P, _I, _I, _I, _P, _P, _P],
    "cnpc_shell_synth_half": [_P, _I, _D, _P, _P, _I, _I, _I, _P, _P, _P, _I, _I, _P, _I],
    "cnpc_contract_zgemm": [_P, _I, _D, _D, _P, _P, _I, _I, _P, _P, _I, _I, _P, _I, _P, _P],
    "cnpc_contract_zgemm_noalloc": [_P, _I, _D, _D, _P, _P, _I, _I, _P, _P, _I, _I, _P, _I, _P, _P],
    "cnpc_contract_zgemm_supported": [_P, _I, _I, _I, _I, ctypes.POINTER(_I)],
    "cnpc_contract_pairs": [_P, _I, _L, _P, _P, _P, _I, _P, _P, _P],
    "cnpc_contract_gemm": [_P, _I, _L, _P, _P, _I, _P, _I, _I, _F, _F, _I, _P],
    "cnpc_absmax_f32": [_P, _P, _L, _P],
    "cnpc_zero": [_P, _P, _L],
    "cnpc_sync": [_P],
    "cnpc_fft_c2c": [_P, _I, _I, _I, _I, _P, _P],
    "cnpc_full_spectrum": [_P, _I, _D, _I, _P, _P, _I, _I, _I, _D, _P],
    "cnpc_xweight": [_P, _I, _D, _I, _P, _P, _P, _P],
    "cnpc_kcombine": [_P, _I, _D, _I, _P, _P, _I, _I, _I, _I, _D, _D, _P, _D, _D, _I, _P],
    "cnpc_full_bin": [_P, _I, _D, _I, _P, _P, _I, _P, _I, _P],
    "cnpc_triple_sum": [_P, _I, _L, _P, _P, _P, _P],
    "cnpc_split_complex": [_P, _I, _L, _P, _P, _P],
    "cnpc_resample_full": [_P, _I, _I, _I, _P, _I, _P],
    "cnpc_reflect_conj": [_P, _I, _I, _P, _D, _D, _P],
    "cnpc_particle_sums": [_P, _L, _P, _P, _P, _P, _I, _P, _P],
    "cnpc_pair_weights": [_P, _L, _P, _P, _I, _D, _P],
    "cnpc_radecz_to_xyz": [_P, _L, _P, _P, _P, _D, _P],
    "cnpc_radecz_to_xyz_flrw": [_P, _L, _P, _P, _P, _P, _D, _P],
}

_lib = None


HIP_ERROR_OUT_OF_MEMORY = 2  # hipErrorOutOfMemory


class CnpcError(RuntimeError):
    """A failed library call; ``status`` is the call's nonzero return code
    (a hipError_t value)."""

    def __init__(self, msg, status=None):
        super().__init__(msg)
        self.status = status


def is_out_of_memory(exc):
    """True for a device allocation failure: a library call that returned
    hipErrorOutOfMemory, or torch's out-of-memory error."""
    if isinstance(exc, CnpcError):
        return exc.status == HIP_ERROR_OUT_OF_MEMORY
    oom = getattr(__import__("torch"), "OutOfMemoryError", None)
    return oom is not None and isinstance(exc, oom)


def load(path=LIB_PATH):
    """Load the shared library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise CnpcError(
            f"{path} not found: build it with `python -m cosmonpc_amd.build_lib` "
            "(the MI355X path has no CPU fallback)")
    lib = ctypes.CDLL(path)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    lib.cnpc_last_error.argtypes = []
    lib.cnpc_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def check(rc, what):
    if rc != 0:
        msg = _lib.cnpc_last_error().decode(errors="replace") if _lib is not None else ""
        raise CnpcError(f"{what} failed (status {rc}): {msg}", rc)
