"""ctypes binding of libcsfunnel.so (declared in include/csfunnel.h).

The library is built in-tree by ``csfunnel/build.py``; importing this module
never falls back to anything else: a missing or unloadable library raises.
"""
from __future__ import annotations

import ctypes as C
import os

NG, NR, NTOT = 23, 6, 29

CS_OK = 0
ERRORS = {
    -1: "BadModel", -2: "NeedTwoElites", -3: "BadArgs", -4: "IO", -5: "BadLine",
    -6: "Parse", -7: "KeepZero", -8: "Capacity", -9: "HIP", -10: "OutOfMemory",
    -11: "Internal", -12: "NoDevice",
}


class CsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code
        self.name = ERRORS.get(code, str(code))


class Genome(C.Structure):
    _fields_ = [("id", C.c_uint64), ("gen", C.c_uint32), ("tribe", C.c_uint32),
                ("pa", C.c_uint64), ("pb", C.c_uint64),
                ("g", C.c_float * NG), ("reg", C.c_float * NR)]


class Model(C.Structure):
    _fields_ = [("w0", C.c_float), ("lin", C.c_float * NTOT), ("quad", C.c_float * 256)]


class ScreenCfg(C.Structure):
    _fields_ = [("lambda_", C.c_uint64), ("keep", C.c_uint64), ("seed", C.c_uint64),
                ("threads", C.c_uint64), ("sigma", C.c_float), ("jump_p", C.c_float),
                ("imm_p", C.c_float), ("_pad", C.c_float)]


class Survivor(C.Structure):
    _fields_ = [("score", C.c_float), ("g", C.c_float * NTOT), ("pa", C.c_uint64),
                ("pb", C.c_uint64), ("tribe", C.c_uint32), ("stream", C.c_uint32),
                ("index", C.c_uint64), ("start", C.c_uint64)]


class ScreenStats(C.Structure):
    _fields_ = [("candidates", C.c_uint64), ("positions", C.c_uint64), ("blocks", C.c_uint64),
                ("records", C.c_uint64), ("levels", C.c_uint32), ("retries", C.c_uint32),
                ("ms_map", C.c_double), ("ms_scan", C.c_double), ("ms_score", C.c_double),
                ("ms_replay", C.c_double), ("ms_total", C.c_double), ("offset_path", C.c_uint32),
                ("walk_fixes", C.c_uint32)]


LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libcsfunnel.so")

# every symbol of include/csfunnel.h, with (restype, argtypes)
SIGNATURES = {
    "cs_version": (C.c_int, []),
    "cs_last_error": (C.c_char_p, []),
    "cs_parse_u64": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(C.c_uint64)]),
    "cs_parse_f32": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(C.c_float)]),
    "cs_parse_pop_line": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(Genome)]),
    "cs_model_from_compact": (C.c_int, [C.POINTER(C.c_float), C.c_int, C.POINTER(Model)]),
    "cs_load_model": (C.c_int, [C.c_char_p, C.POINTER(Model)]),
    "cs_load_elites": (C.c_int, [C.c_char_p, C.POINTER(Genome), C.c_int, C.POINTER(C.c_int)]),
    "cs_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "cs_ctx_destroy": (C.c_int, [C.c_void_p]),
    "cs_screen": (C.c_int, [C.c_void_p, C.POINTER(Model), C.POINTER(Genome), C.c_int,
                            C.POINTER(ScreenCfg), C.POINTER(Survivor), C.c_uint64,
                            C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "cs_screen_noalloc": (C.c_int, [C.c_void_p, C.POINTER(Model), C.POINTER(Genome), C.c_int,
                                    C.POINTER(ScreenCfg), C.POINTER(Survivor), C.c_uint64,
                                    C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "cs_get_stats": (C.c_int, [C.c_void_p, C.POINTER(ScreenStats)]),
    "cs_set_record_cap": (C.c_int, [C.c_void_p, C.c_uint32]),
    "cs_set_offset_mode": (C.c_int, [C.c_void_p, C.c_int]),
    "cs_mlp_create": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    "cs_mlp_destroy": (C.c_int, [C.c_void_p]),
    "cs_mlp_get_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "cs_mlp_set_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "cs_mlp_set_fp8_scale": (C.c_int, [C.c_void_p, C.c_float]),
    "cs_mlp_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_uint64, C.c_void_p, C.POINTER(C.c_float)]),
    "cs_mlp_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_float),
                              C.POINTER(C.c_float)]),
    "cs_mlp_train_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_float, C.c_float,
                                    C.POINTER(C.c_float)]),
    "cs_screen_ex": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Genome), C.c_int, C.POINTER(ScreenCfg),
                               C.c_void_p, C.POINTER(Survivor), C.c_uint64, C.POINTER(C.c_uint64),
                               C.POINTER(C.c_uint64)]),
    "cs_ridge_gram_f64": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_uint64, C.c_void_p,
                                    C.c_void_p, C.c_void_p]),
    "cs_ridge_solve_f64": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_void_p,
                                     C.c_void_p]),
    "cs_ridge_loss_grad_f64": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_uint64, C.c_void_p,
                                         C.c_void_p, C.c_double, C.c_void_p, C.c_void_p]),
    "cs_format_gene": (C.c_int, [C.c_float, C.c_char_p, C.c_size_t]),
    "cs_write_promoted": (C.c_int, [C.c_char_p, C.POINTER(Survivor), C.c_uint64]),
    "cs_write_promoted_v6": (C.c_int, [C.c_char_p, C.POINTER(Survivor), C.c_uint64]),
    "cs_write_count": (C.c_int, [C.c_char_p, C.c_uint64]),
    "cs_valu_rate_probe": (C.c_int, [C.c_int, C.c_int, C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_uint32)]),
}

_lib = None


def _bind_one_hip_runtime() -> None:
    """Make every user of the HIP runtime in this process share one copy.

    torch's ROCm wheel links `libamdhip64.so` / `libhsa-runtime64.so` by their
    unversioned names from its own lib directory, this library links
    `libamdhip64.so.7`.  Loading the runtime under the unversioned names first
    lets both bind to the same copy whichever is imported first (if torch is
    already loaded these calls return its copy)."""
    for name in ("libhsa-runtime64.so", "libamdhip64.so"):
        try:
            C.CDLL(name, mode=C.RTLD_GLOBAL)
        except OSError:
            pass


def lib() -> C.CDLL:
    """Load libcsfunnel.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run csfunnel/build.py")
        _bind_one_hip_runtime()
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> int:
    if rc < 0:
        raise CsError(rc, lib().cs_last_error().decode(errors="replace"))
    return rc
