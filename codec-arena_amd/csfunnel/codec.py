"""The codec label source: genome -> (PSNR, bpp, ...) over the synthetic corpus.

ctypes binding of libcscodec.so (include/cscodec.h), the CPU restatement of
codec-arena's encoder/decoder, and the corpus-level aggregation of `arena
eval` (src/main.zig:283-344): PSNR, chroma PSNR and MS-SSIM are f32 means over
clips, bpp is total bytes * 8 / total samples (luma + chroma), dec_mops is the
f32 sum of per-clip decoder mops, msssim_db is the mean of per-clip
-10 log10(1 - MS-SSIM).  The label the learned surrogate is trained on is
PSNR - 14 bpp (brain/config.json:2, brain/evolve.py:197-205 with the PSNR
quality key).

Genomes are evaluated in parallel on host threads (ctypes releases the GIL).
"""
from __future__ import annotations

import ctypes as C
import os
from concurrent.futures import ThreadPoolExecutor
from typing import Dict, Iterable, List, Optional, Sequence

import numpy as np

from . import corpus as _corpus
from ._abi import NG, NR, Genome

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libcscodec.so")
LAMBDA = 14.0  # brain/config.json:2

FAST_SAD, CLOSED_LOOP, MSSSIM = 1, 2, 4
ERRORS = {-20: "Unsupported", -21: "Args", -22: "Overflow", -23: "Corrupt", -24: "Desync"}


class ClipStats(C.Structure):
    _fields_ = [("bytes", C.c_uint64), ("psnr_y", C.c_float), ("psnr_u", C.c_float), ("psnr_v", C.c_float),
                ("msssim_y", C.c_float), ("enc_mops", C.c_float), ("dec_mops", C.c_float),
                ("dct_calls", C.c_uint64), ("idct_calls", C.c_uint64), ("closed_loop", C.c_int32),
                ("enc_ms", C.c_float), ("sb_unified", C.c_uint32 * 3)]


class CodecError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{ERRORS.get(code, code)}: {what}")
        self.code = code


SIGNATURES = {
    "csc_genome_supported": (C.c_int, [C.POINTER(Genome)]),
    "csc_eval_clip": (C.c_int, [C.c_char_p, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(Genome), C.c_uint32,
                                C.POINTER(ClipStats)]),
    "csc_eval_clip_ex": (C.c_int, [C.c_char_p, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(Genome), C.c_uint32,
                                   C.c_float, C.c_void_p, C.POINTER(ClipStats)]),
    "csc_plane_msssim": (C.c_int, [C.c_char_p, C.c_char_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_float)]),
    "csc_psnr_plane": (C.c_float, [C.c_char_p, C.c_char_p, C.c_uint64]),
    "csc_svarint_roundtrip": (C.c_int64, [C.c_int64]),
    "csc_intra_predict": (C.c_int, [C.c_int, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_int,
                                    C.c_char_p]),
    "csc_interp_q": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                               C.c_char_p]),
    "csc_deblock_pair": (C.c_int, [C.POINTER(C.c_uint8), C.POINTER(C.c_uint8), C.c_uint8, C.c_uint8, C.c_float,
                                   C.c_float]),
    "csc_rc_roundtrip": (C.c_int64, [C.c_char_p, C.c_char_p, C.c_uint64, C.c_char_p, C.c_uint64, C.c_char_p]),
}

_lib = None


def lib() -> C.CDLL:
    """Load libcscodec.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run csfunnel/build.py")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def genome(genes: Sequence[float], gid: int = 0) -> Genome:
    """A Genome from a flat vector: 16 legacy effectors (rest 0 and
    regulators 0.5, the v1 migration of src/genome.zig:116-188), 23
    effectors (regulators 0.5), 28 (v5: 22 effectors + 6 regulators) or 29
    (23 + 6) values."""
    v = [float(x) for x in genes]
    g = Genome()
    g.id = gid
    if len(v) in (16, 23):
        eff, reg = v, [0.5] * NR
    elif len(v) == 28:
        eff, reg = v[:22], v[22:]
    elif len(v) == NG + NR:
        eff, reg = v[:NG], v[NG:]
    else:
        raise ValueError(f"genome vector of length {len(v)}")
    for k in range(NG):
        g.g[k] = eff[k] if k < len(eff) else 0.0
    for k in range(NR):
        g.reg[k] = reg[k]
    return g


def supported(g: Genome) -> bool:
    return bool(lib().csc_genome_supported(C.byref(g)))


def eval_clip(data: bytes, g: Genome, w: int = _corpus.W, h: int = _corpus.H, frames: int = _corpus.FRAMES,
              flags: int = FAST_SAD, qscale: float = 1.0, recon: Optional[bytearray] = None) -> ClipStats:
    """evalClip of one clip; `recon`, when given, must hold frames*w*h*3/2
    bytes and receives the decoded clip."""
    st = ClipStats()
    if recon is None and qscale == 1.0:
        rc = lib().csc_eval_clip(data, w, h, frames, C.byref(g), flags, C.byref(st))
    else:
        buf = None
        if recon is not None:
            if len(recon) < frames * w * h * 3 // 2:
                raise ValueError("recon buffer smaller than the clip")
            buf = (C.c_uint8 * len(recon)).from_buffer(recon)
        rc = lib().csc_eval_clip_ex(data, w, h, frames, C.byref(g), flags, qscale, buf, C.byref(st))
    if rc < 0:
        raise CodecError(rc, f"genome {g.id}")
    return st


def _aggregate(stats: List[ClipStats], clip_px: List[int]) -> Dict[str, float]:
    f32 = np.float32
    n = f32(len(stats))
    acc = {k: f32(0) for k in ("psnr", "psnr_u", "psnr_v", "msssim", "ms_db", "dec", "enc")}
    for st in stats:
        acc["psnr"] = f32(acc["psnr"] + f32(st.psnr_y))
        acc["psnr_u"] = f32(acc["psnr_u"] + f32(st.psnr_u))
        acc["psnr_v"] = f32(acc["psnr_v"] + f32(st.psnr_v))
        acc["msssim"] = f32(acc["msssim"] + f32(st.msssim_y))
        m = min(max(f32(st.msssim_y), f32(1e-6)), f32(1.0 - 1e-9))
        acc["ms_db"] = f32(acc["ms_db"] + f32(f32(-10.0) * np.log10(f32(f32(1.0) - m))))
        acc["dec"] = f32(acc["dec"] + f32(st.dec_mops))
        acc["enc"] = f32(acc["enc"] + f32(st.enc_ms))
    tot_bytes = sum(int(st.bytes) for st in stats)
    bpp = float(tot_bytes) * 8.0 / float(sum(clip_px))
    psnr = float(f32(acc["psnr"] / n))
    return {"psnr": psnr, "bpp": bpp, "dec_mops": float(acc["dec"]), "psnr_u": float(f32(acc["psnr_u"] / n)),
            "psnr_v": float(f32(acc["psnr_v"] / n)), "msssim": float(f32(acc["msssim"] / n)),
            "msssim_db": float(f32(acc["ms_db"] / n)), "enc_ms10": float(f32(acc["enc"] * f32(10.0))),
            "bytes": tot_bytes, "label": psnr - LAMBDA * bpp}


class Evaluator:
    """`arena eval` over a set of corpus clips (default: all ten)."""

    def __init__(self, names: Iterable[str] = _corpus.NAMES, corpus_dir: Optional[str] = None,
                 threads: Optional[int] = None, fast_sad: bool = True, msssim: bool = False):
        self.names = tuple(names)
        data = _corpus.load(corpus_dir, self.names)
        self.clips = [data[n] for n in self.names]
        self.flags = (FAST_SAD if fast_sad else 0) | (MSSSIM if msssim else 0)
        self.threads = threads or min(16, os.cpu_count() or 1)
        self.px = [_corpus.FRAMES * _corpus.W * _corpus.H * 3 // 2] * len(self.clips)
        lib()

    def metrics(self, g: Genome) -> Dict[str, float]:
        return _aggregate([eval_clip(c, g, flags=self.flags) for c in self.clips], self.px)

    def evaluate(self, genomes: Sequence, ids: Optional[Sequence[int]] = None) -> List[Dict[str, float]]:
        """Metrics for every genome (Genome structs or flat gene vectors),
        clip jobs spread over the host threads."""
        gs = [g if isinstance(g, Genome) else genome(g, i if ids is None else ids[i]) for i, g in enumerate(genomes)]
        jobs = [(i, j) for i in range(len(gs)) for j in range(len(self.clips))]
        with ThreadPoolExecutor(self.threads) as ex:
            res = list(ex.map(lambda ij: eval_clip(self.clips[ij[1]], gs[ij[0]], flags=self.flags), jobs))
        out, k = [], 0
        for _ in gs:
            out.append(_aggregate(res[k:k + len(self.clips)], self.px))
            k += len(self.clips)
        return out

    def labels(self, genomes: Sequence) -> np.ndarray:
        """PSNR - 14 bpp per genome."""
        return np.array([m["label"] for m in self.evaluate(genomes)], dtype=np.float64)
