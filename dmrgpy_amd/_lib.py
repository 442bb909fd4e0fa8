"""ctypes binding of libdmrg_hip.so (C ABI declared in include/dmrg_hip.h).

There is no fallback: if the library is missing or cannot be loaded, every
entry into the engine raises.
"""

import ctypes
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libdmrg_hip.so")

c_int = ctypes.c_int
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p
c_ll = ctypes.c_longlong
P_double = ctypes.POINTER(ctypes.c_double)
P_int = ctypes.POINTER(ctypes.c_int)
P_ll = ctypes.POINTER(ctypes.c_longlong)


class HeffDesc(ctypes.Structure):
    _fields_ = [("nsite", c_int), ("field", c_int), ("prec", c_int),
                ("Dl", c_int), ("Dr", c_int), ("d1", c_int), ("d2", c_int),
                ("wl", c_int), ("wm", c_int), ("wr", c_int),
                ("L", c_void_p), ("R", c_void_p),
                ("W1", c_void_p), ("W2", c_void_p), ("w_complex", c_int),
                ("ldL", c_int), ("ldR", c_int),
                ("nproj", c_int), ("proj", c_void_p), ("proj_weight", c_double)]


class EnvDesc(ctypes.Structure):
    _fields_ = [("dir", c_int), ("field", c_int), ("prec", c_int),
                ("Dl", c_int), ("Dr", c_int), ("Dlb", c_int), ("Drb", c_int),
                ("d", c_int), ("wl", c_int), ("wr", c_int),
                ("ket", c_void_p), ("bra", c_void_p), ("W", c_void_p), ("w_complex", c_int),
                ("env_in", c_void_p), ("env_out", c_void_p),
                ("ld_in", c_int), ("ld_out", c_int)]


# name -> (restype, argtypes); every symbol include/dmrg_hip.h declares
SIGNATURES = {
    "dmrg_version": (c_int, []),
    "dmrg_last_error": (ctypes.c_char_p, []),
    "dmrg_ctx_create": (c_int, [c_int, c_void_p, ctypes.POINTER(c_void_p)]),
    "dmrg_ctx_destroy": (c_int, [c_void_p]),
    "dmrg_sync": (c_int, [c_void_p]),
    "dmrg_heff_apply": (c_int, [c_void_p, ctypes.POINTER(HeffDesc), c_void_p, c_void_p]),
    "dmrg_heff_apply_noalloc": (c_int, [c_void_p, ctypes.POINTER(HeffDesc), c_void_p, c_void_p]),
    "dmrg_rayleigh": (c_int, [c_void_p, ctypes.POINTER(HeffDesc), c_void_p, P_double]),
    "dmrg_env_update": (c_int, [c_void_p, ctypes.POINTER(EnvDesc)]),
    "dmrg_lanczos_ground": (c_int, [c_void_p, ctypes.POINTER(HeffDesc), c_void_p, c_int, c_double,
                                    c_double, P_double, P_int]),
    "dmrg_krylov_expm": (c_int, [c_void_p, ctypes.POINTER(HeffDesc), c_void_p, c_double, c_double,
                                 c_int, c_double, c_double, P_int]),
    "dmrg_svd_split": (c_int, [c_void_p, c_int, c_void_p, c_int, c_int, c_double, c_int, c_int, c_int,
                               c_void_p, c_void_p, c_int, P_int, P_double, P_double]),
    "dmrg_svd_split_guided": (c_int, [c_void_p, c_int, c_void_p, c_int, c_int, c_double, c_int, c_int, c_int,
                                      c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_int, P_int, P_double,
                                      P_double, P_int, P_double]),
    "dmrg_set_eigensolver": (c_int, [c_void_p, c_int]),
    "dmrg_sym_eig": (c_int, [c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p]),
    "dmrg_herm_eig": (c_int, [c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p]),
    "dmrg_mpo_grad": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int,
                              c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_void_p]),
    "dmrg_gemm": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                          c_double, c_void_p, c_int, c_ll, c_ll, c_void_p, c_int, c_ll, c_ll,
                          c_double, c_void_p, c_int, c_ll, c_ll, c_int]),
    "dmrg_prof_enable": (c_int, [c_void_p, c_int]),
    "dmrg_prof_read": (c_int, [c_void_p, P_double, P_double, P_ll]),
    "dmrg_prof_read_bytes": (c_int, [c_void_p, P_double]),
    "dmrg_matvec_count": (c_ll, [c_void_p]),
}

_lib = None


def load():
    """The loaded library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libdmrg_hip.so is not built ({}); run `python -m dmrgpy_amd.build`".format(LIB_PATH))
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status, what=""):
    if status != 0:
        msg = load().dmrg_last_error()
        raise RuntimeError("{} failed ({}): {}".format(what, status, msg.decode() if msg else "?"))
