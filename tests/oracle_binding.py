"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is the parity checker only; nothing in the product imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
LIB = os.path.join(ORACLE_DIR, "liboracle.so")

NTOT = 29


class OrGenome(C.Structure):
    _fields_ = [("id", C.c_uint64), ("gen", C.c_uint32), ("tribe", C.c_uint32),
                ("pa", C.c_uint64), ("pb", C.c_uint64),
                ("g", C.c_float * 23), ("reg", C.c_float * 6)]


class OrModel(C.Structure):
    _fields_ = [("w0", C.c_float), ("lin", C.c_float * 29), ("quad", C.c_float * 256)]


class OrCfg(C.Structure):
    _fields_ = [("lambda_", C.c_uint64), ("keep", C.c_uint64), ("seed", C.c_uint64),
                ("threads", C.c_uint64), ("sigma", C.c_float), ("jump_p", C.c_float),
                ("imm_p", C.c_float), ("_pad", C.c_float)]


SLOT_DTYPE = np.dtype([("score", "<f4"), ("g", "<f4", (29,)), ("pa", "<u8"), ("pb", "<u8"),
                       ("tribe", "<u4"), ("stream", "<u4"), ("index", "<u8"), ("start", "<u8")])


def build_oracle() -> str:
    if not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    return LIB


class Oracle:
    def __init__(self):
        self.L = C.CDLL(build_oracle())
        L = self.L
        L.or_splitmix_fill.argtypes = [C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64)]
        L.or_stream_seed.restype = C.c_uint64
        L.or_stream_seed.argtypes = [C.c_uint64, C.c_uint64]
        L.or_parse_line.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(OrGenome)]
        L.or_model_from_compact.argtypes = [C.POINTER(C.c_float), C.c_int, C.POINTER(OrModel)]
        L.or_load_model.argtypes = [C.c_char_p, C.POINTER(OrModel)]
        L.or_load_elites.argtypes = [C.c_char_p, C.POINTER(OrGenome), C.c_int, C.POINTER(C.c_int)]
        L.or_score.restype = C.c_float
        L.or_score.argtypes = [C.POINTER(OrModel), C.POINTER(C.c_float)]
        L.or_score_simd.restype = C.c_float
        L.or_score_simd.argtypes = [C.POINTER(OrModel), C.POINTER(C.c_float)]
        L.or_clamp01.restype = C.c_float
        L.or_clamp01.argtypes = [C.c_float]
        L.or_synth.argtypes = [C.POINTER(OrGenome), C.c_int, C.c_uint64, C.c_uint64, C.c_float, C.c_float,
                               C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_uint32),
                               C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.or_screen.argtypes = [C.POINTER(OrModel), C.POINTER(OrGenome), C.c_int, C.POINTER(OrCfg), C.c_int,
                                C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.or_screen_simd.argtypes = L.or_screen.argtypes
        L.or_screen_chunked.argtypes = [C.POINTER(OrModel), C.POINTER(OrGenome), C.c_int, C.POINTER(OrCfg), C.c_int,
                                        C.c_int, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.or_fmt_d4.argtypes = [C.c_float, C.c_char_p, C.c_size_t]
        L.or_write_promoted.argtypes = [C.c_char_p, C.c_void_p, C.c_uint64]
        L.or_write_count.argtypes = [C.c_char_p, C.c_uint64]

    # --- primitives
    def splitmix(self, seed: int, n: int) -> np.ndarray:
        out = np.zeros(n, dtype=np.uint64)
        self.L.or_splitmix_fill(seed, n, out.ctypes.data_as(C.POINTER(C.c_uint64)))
        return out

    def stream_seed(self, seed: int, t: int) -> int:
        return self.L.or_stream_seed(seed, t)

    def parse_line(self, line: str):
        g = OrGenome()
        b = line.encode()
        rc = self.L.or_parse_line(b, len(b), C.byref(g))
        return rc, g

    def model_from_compact(self, w) -> OrModel:
        arr = np.ascontiguousarray(np.asarray(w, dtype=np.float32))
        m = OrModel()
        rc = self.L.or_model_from_compact(arr.ctypes.data_as(C.POINTER(C.c_float)), int(arr.size), C.byref(m))
        if rc != 0:
            raise ValueError(f"or_model_from_compact rc={rc}")
        return m

    def load_model(self, path: str) -> OrModel:
        m = OrModel()
        rc = self.L.or_load_model(path.encode(), C.byref(m))
        if rc != 0:
            raise ValueError(f"or_load_model rc={rc}")
        return m

    def load_elites(self, path: str):
        arr = (OrGenome * 200000)()
        n = C.c_int()
        rc = self.L.or_load_elites(path.encode(), arr, 200000, C.byref(n))
        if rc != 0:
            raise ValueError(f"or_load_elites rc={rc}")
        return (OrGenome * n.value)(*arr[: n.value])

    def score(self, m: OrModel, g) -> float:
        arr = np.ascontiguousarray(np.asarray(g, dtype=np.float32))
        return self.L.or_score(C.byref(m), arr.ctypes.data_as(C.POINTER(C.c_float)))

    def score_simd(self, m: OrModel, g) -> float:
        arr = np.ascontiguousarray(np.asarray(g, dtype=np.float32))
        return self.L.or_score_simd(C.byref(m), arr.ctypes.data_as(C.POINTER(C.c_float)))

    def synth(self, elites, stream_seed: int, start: int, sigma: float, jump_p: float, imm_p: float):
        g = (C.c_float * NTOT)()
        L = C.c_uint32()
        ia = C.c_uint32()
        ib = C.c_uint32()
        self.L.or_synth(elites, len(elites), stream_seed, start, sigma, jump_p, imm_p, g, C.byref(L),
                        C.byref(ia), C.byref(ib))
        return np.array(g[:], dtype=np.float32), L.value, ia.value, ib.value

    def screen(self, m: OrModel, elites, lam: int, keep: int, seed: int, threads: int = 4, sigma: float = 0.12,
               jump_p: float = 0.02, imm_p: float = 0.01, workers: int = 0, simd: bool = False):
        """or_screen, or with simd=True or_screen_simd (the 16-wide score form;
        identical results, the form the CPU baseline times)."""
        cfg = OrCfg(lam, keep, seed, threads, sigma, jump_p, imm_p, 0.0)
        cap = max(1, keep)
        out = np.zeros(cap, dtype=SLOT_DTYPE)
        n = C.c_uint64()
        scr = C.c_uint64()
        w = workers or min(os.cpu_count() or 1, 16)
        fn = self.L.or_screen_simd if simd else self.L.or_screen
        rc = fn(C.byref(m), elites, len(elites), C.byref(cfg), w, out.ctypes.data, cap, C.byref(n), C.byref(scr))
        if rc != 0:
            raise ValueError(f"or_screen rc={rc}")
        return out[: n.value].copy(), scr.value

    def screen_chunked(self, m: OrModel, elites, lam: int, keep: int, seed: int, threads: int = 4,
                       sigma: float = 0.12, jump_p: float = 0.02, imm_p: float = 0.01, workers: int = 0,
                       chunks: int = 64):
        """or_screen_chunked: the same survivors as screen(), each stream cut
        into `chunks` pieces run on up to `workers` host threads."""
        cfg = OrCfg(lam, keep, seed, threads, sigma, jump_p, imm_p, 0.0)
        cap = max(1, keep)
        out = np.zeros(cap, dtype=SLOT_DTYPE)
        n = C.c_uint64()
        scr = C.c_uint64()
        w = workers or min(os.cpu_count() or 1, 16)
        rc = self.L.or_screen_chunked(C.byref(m), elites, len(elites), C.byref(cfg), w, chunks, out.ctypes.data, cap,
                                      C.byref(n), C.byref(scr))
        if rc != 0:
            raise ValueError(f"or_screen_chunked rc={rc}")
        return out[: n.value].copy(), scr.value

    def fmt(self, v: float) -> str:
        buf = C.create_string_buffer(64)
        self.L.or_fmt_d4(C.c_float(v), buf, 64)
        return buf.value.decode()

    def write_promoted(self, path: str, slots: np.ndarray):
        arr = np.ascontiguousarray(slots, dtype=SLOT_DTYPE)
        self.L.or_write_promoted(path.encode(), arr.ctypes.data, len(arr))

    def write_count(self, path: str, total: int):
        self.L.or_write_count(path.encode(), total)
