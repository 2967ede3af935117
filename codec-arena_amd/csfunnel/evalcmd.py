"""The codec subcommands of `arena` on the codec label source: `eval`, `psnr`
and `msssimcheck` (src/main.zig:32-56, 58-95, 103-125, 220-356).

`eval` reads the corpus spec (DIR/corpus.txt, lines "name w h frames fps"),
the optional clip subset, and the population file, evaluates every parsed
genome on every clip through libcscodec.so, and writes one CSV line per genome
that evaluated on all clips:
    id,psnr,bpp,enc_ms_x10,dec_mops,psnr_u,psnr_v,msssim,msssim_db
with the reference's aggregation (f32 means over clips, bpp from total bytes
over total luma + chroma samples, msssim_db the mean of per-clip
-10 log10(1 - M)) and its decimal formatting.  Genome x clip jobs run on host
threads; lines keep the population order.  Genomes that evaluate with an error
are counted as bad and reported on stderr, as the reference does.
"""
from __future__ import annotations

import math
import os
import sys
from concurrent.futures import ThreadPoolExecutor
from typing import List, Optional, Tuple

import numpy as np

from . import codec
from .screen import CsError, parse_f32, parse_line, parse_u64


class CmdError(Exception):
    """A Zig error returned from main: exit code 1, "error: <name>" on stderr."""


def zig_decimal(v: float, prec: int, bits: int = 32) -> str:
    """Zig std.fmt `{d:.prec}` of an f32 (bits=32) or f64 value: the shortest
    round-trip digits of the value, cut at `prec` decimal places and rounded
    half-up on the first removed digit, printed with zero padding."""
    if math.isnan(v):
        return "nan"
    if math.isinf(v):
        return "inf" if v > 0 else "-inf"
    x = np.float32(v) if bits == 32 else np.float64(v)
    out = "-" if np.signbit(x) else ""
    a = abs(x)
    if a == 0:
        dg, exp10 = "0", 0
    else:
        mant, e = np.format_float_scientific(a, unique=True, trim="-").split("e")
        dg = mant.replace(".", "")
        exp10 = int(e) + 1 - len(dg)  # value = int(dg) * 10^exp10
        while len(dg) > 1 and dg[-1] == "0":
            dg = dg[:-1]
            exp10 += 1
    ol = len(dg)
    if exp10 > 0:
        keep = (ol - 1) + prec + exp10
    else:
        keep = prec + ol + exp10 if prec + ol > -exp10 else 0
    if keep < ol:
        first_removed = dg[keep]
        kept = list(dg[:keep])
        exp10 += ol - keep
        if first_removed >= "5":
            i = len(kept) - 1
            while i >= 0 and kept[i] == "9":
                kept[i] = "0"
                i -= 1
            if i >= 0:
                kept[i] = chr(ord(kept[i]) + 1)
            else:
                kept.insert(0, "1")
        dg = "".join(kept).lstrip("0") or "0"
    ol = len(dg)
    point = exp10 + ol
    if point <= 0:
        s = "0." + "0" * (-point) + dg
        s += "0" * max(0, prec - (-point + ol))
    elif point >= ol:
        s = dg + "0" * (point - ol) + "." + "0" * prec
    else:
        s = dg[:point] + "." + dg[point:]
        s += "0" * max(0, prec - (ol - point))
    return out + s


def _u32(tok: str) -> int:
    try:
        v = parse_u64(tok)
    except CsError as e:
        raise CmdError("InvalidCharacter") from e
    if v > 0xFFFFFFFF:
        raise CmdError("Overflow")
    return v


def _read(path: str, limit: int) -> bytes:
    """readFileAlloc with a size limit (error.FileTooBig beyond it)."""
    try:
        with open(path, "rb") as f:
            data = f.read(limit + 1)
    except FileNotFoundError as e:
        raise CmdError("FileNotFound") from e
    except OSError as e:
        raise CmdError("AccessDenied") from e
    if len(data) > limit:
        raise CmdError("FileTooBig")
    return data


def load_clip(path: str, w: int, h: int, frames: int) -> bytes:
    """yuv.loadClip (src/yuv.zig:34-62): at least frames * w * h * 3/2 bytes."""
    data = _read(path, 1 << 28)
    ysz = w * h
    total = frames * (ysz + 2 * (ysz // 4))
    if len(data) < total:
        raise CmdError("CorruptClip")
    return data[:total]


def _lines(txt: bytes):
    for raw in txt.split(b"\n"):
        line = raw.strip(b" \t\r")
        if line and not line.startswith(b"#"):
            yield line.decode("utf-8", "replace")


def load_corpus(corpus_dir: str, clips_path: str = "") -> List[Tuple[str, int, int, int, bytes]]:
    """The corpus spec and clip subset of evalAll (src/main.zig:231-275)."""
    clips = []
    for line in _lines(_read(os.path.join(corpus_dir, "corpus.txt"), 1 << 20)):
        t = line.replace("\t", " ").split()
        if not t:
            continue
        if len(t) < 5:
            raise CmdError("BadSpec")
        w, h, fr = _u32(t[1]), _u32(t[2]), _u32(t[3])
        _u32(t[4])
        clips.append((t[0], w, h, fr, load_clip(os.path.join(corpus_dir, t[0] + ".yuv"), w, h, fr)))
    if not clips:
        raise CmdError("NoClips")
    if clips_path:
        keep = []
        for name in _lines(_read(clips_path, 1 << 16)):
            keep.extend(c for c in clips if c[0] == name)
        if not keep:
            raise CmdError("NoClipsMatched")
        clips = keep
    return clips


def _genome_line(g, stats: List[codec.ClipStats], clips) -> str:
    f32 = np.float32
    n = f32(len(clips))
    acc = dict.fromkeys(("psnr", "u", "v", "ms", "db", "enc", "dec"), f32(0))
    tot_bytes = tot_px = 0
    for st, c in zip(stats, clips):
        tot_bytes += int(st.bytes)
        acc["psnr"] = f32(acc["psnr"] + f32(st.psnr_y))
        acc["u"] = f32(acc["u"] + f32(st.psnr_u))
        acc["v"] = f32(acc["v"] + f32(st.psnr_v))
        acc["ms"] = f32(acc["ms"] + f32(st.msssim_y))
        m = min(max(f32(st.msssim_y), f32(1e-6)), f32(1.0 - 1e-9))
        acc["db"] = f32(acc["db"] + f32(f32(-10.0) * np.log10(f32(f32(1.0) - m))))
        acc["enc"] = f32(acc["enc"] + f32(st.enc_ms))
        acc["dec"] = f32(acc["dec"] + f32(st.dec_mops))
        tot_px += c[3] * (c[1] * c[2]) * 3 // 2
    bpp = float(tot_bytes) * 8.0 / float(tot_px)
    return ",".join([str(g.id), zig_decimal(acc["psnr"] / n, 2), zig_decimal(bpp, 4, 64),
                     zig_decimal(acc["enc"] * f32(10.0), 1), zig_decimal(acc["dec"], 2),
                     zig_decimal(acc["u"] / n, 2), zig_decimal(acc["v"] / n, 2),
                     zig_decimal(acc["ms"] / n, 5), zig_decimal(acc["db"] / n, 4)]) + "\n"


def eval_cmd(argv: List[str], threads: Optional[int] = None) -> int:
    """`arena eval` (src/main.zig:58-95 flag parsing, :220-356 evalAll)."""
    corpus_dir, pop_path, out_path, clips_path = "corpus/raw", "pop/seed.txt", "results.csv", ""
    fast, qscale, dump = True, 1.0, None
    try:
        i = 0
        while i < len(argv):
            a, has = argv[i], i + 1 < len(argv)
            if a == "--corpus" and has:
                i += 1
                corpus_dir = argv[i]
            elif a == "--pop" and has:
                i += 1
                pop_path = argv[i]
            elif a == "--out" and has:
                i += 1
                out_path = argv[i]
            elif a == "--clips" and has:
                i += 1
                clips_path = argv[i]
            elif a == "--slow":
                fast = False
            elif a == "--qscale" and has:
                i += 1
                try:
                    qscale = parse_f32(argv[i])
                except CsError as e:
                    raise CmdError("InvalidCharacter") from e
            elif a == "--dump-recon" and has:
                i += 1
                dump = argv[i]
            i += 1
        clips = load_corpus(corpus_dir, clips_path)
        pop = _read(pop_path, 1 << 24)
        genomes, n_bad = [], 0
        for line in _lines(pop):
            try:
                g = parse_line(line)
            except CsError:
                n_bad += 1
                continue
            if g is not None:
                genomes.append(g)
        flags = (codec.FAST_SAD if fast else 0) | codec.MSSSIM
        codec.lib()

        def job(k):
            gi, ci = divmod(k, len(clips))
            _, w, h, fr, data = clips[ci]
            recon = bytearray(len(data)) if dump else None
            try:
                st = codec.eval_clip(data, genomes[gi], w, h, fr, flags, qscale, recon)
            except codec.CodecError as e:
                return None, codec.ERRORS.get(e.code, str(e.code)), None
            return st, None, recon

        nthreads = threads or int(os.environ.get("CSFUNNEL_EVAL_THREADS", 0)) or min(16, os.cpu_count() or 1)
        n_ok = 0
        # genomes go to the pool a chunk at a time, so the held results (and
        # decoded clips, with --dump-recon) stay bounded
        chunk = 1 if dump else 64
        nc = len(clips)
        with open(out_path, "w") as out, ThreadPoolExecutor(nthreads) as ex:
            for g0 in range(0, len(genomes), chunk):
                g1 = min(len(genomes), g0 + chunk)
                res = list(ex.map(job, range(g0 * nc, g1 * nc)))
                for gi in range(g0, g1):
                    row = res[(gi - g0) * nc:(gi - g0 + 1) * nc]
                    bad = next((k for k, r in enumerate(row) if r[0] is None), None)
                    if dump:
                        os.makedirs(dump, exist_ok=True)
                        for k, r in enumerate(row[:nc if bad is None else bad]):
                            with open(os.path.join(dump, clips[k][0] + ".recon.yuv"), "wb") as f:
                                f.write(r[2])
                    if bad is not None:
                        sys.stderr.write(f"[evalClip error: error.{row[bad][1]} clip={clips[bad][0]}]\n")
                        n_bad += 1
                        continue
                    out.write(_genome_line(genomes[gi], [r[0] for r in row], clips))
                    n_ok += 1
        sys.stderr.write(f"eval: {n_ok} ok, {n_bad} bad, {len(clips)} clips\n")
        return 0
    except CmdError as e:
        sys.stderr.write(f"error: {e}\n")
        return 1


def psnr_cmd(argv: List[str]) -> int:
    """`arena psnr A.yuv B.yuv W H FRAMES` (src/main.zig:32-55): mean Y, U, V
    PSNR over frames, printed "{d:.4} {d:.4} {d:.4}"."""
    try:
        if len(argv) < 5:
            raise CmdError("BadArgs")
        w, h, fr = _u32(argv[2]), _u32(argv[3]), _u32(argv[4])
        a, b = load_clip(argv[0], w, h, fr), load_clip(argv[1], w, h, fr)
        ysz, csz = w * h, (w * h) // 4
        L = codec.lib()
        f32 = np.float32
        acc = [f32(0)] * 3
        for i in range(fr):
            offs = ((i * ysz, ysz), (fr * ysz + i * csz, csz), (fr * (ysz + csz) + i * csz, csz))
            for p, (o, n) in enumerate(offs):
                acc[p] = f32(acc[p] + f32(L.csc_psnr_plane(a[o:o + n], b[o:o + n], n)))
        nf = f32(fr)
        sys.stdout.write(" ".join(zig_decimal(x / nf, 4) for x in acc) + "\n")
        return 0
    except CmdError as e:
        sys.stderr.write(f"error: {e}\n")
        return 1


def msssimcheck_cmd(argv: List[str]) -> int:
    """`arena msssimcheck PAIRS.bin` (src/main.zig:103-125): records of u32 w,
    u32 h, w*h bytes A, w*h bytes B; prints planeMsssim per record as {d:.6}."""
    import ctypes as C
    try:
        if not argv:
            raise CmdError("BadArgs")
        data = _read(argv[0], 1 << 28)
        off = 0
        L = codec.lib()
        while off + 8 <= len(data):
            w = int.from_bytes(data[off:off + 4], "little")
            h = int.from_bytes(data[off + 4:off + 8], "little")
            off += 8
            n = w * h
            if off + 2 * n > len(data):
                break
            v = C.c_float()
            rc = L.csc_plane_msssim(data[off:off + n], data[off + n:off + 2 * n], w, h, C.byref(v))
            off += 2 * n
            if rc < 0:
                raise CmdError(codec.ERRORS.get(rc, str(rc)))
            sys.stdout.write(zig_decimal(v.value, 6) + "\n")
        return 0
    except CmdError as e:
        sys.stderr.write(f"error: {e}\n")
        return 1
