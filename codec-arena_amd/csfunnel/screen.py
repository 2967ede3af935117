"""Host-side mirror of the reference screen interface.

Names and behaviour follow the reference (codec-arena snapshot):
  Model.from_compact / load_model   src/screen.zig:36-95, 115-127
  Cfg (defaults)                    src/screen.zig:129-137
  Screen.run                        src/screen.zig:257-343 (device path)
  parse_line / load_elites          src/genome.zig:116-188, src/main.zig:201-214
  screen_cmd                        src/main.zig:156-218 (the `arena screen` CLI)
All computation happens in libcsfunnel (HIP kernels + C++ host code); this
module only marshals arguments.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
import sys
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _abi
from ._abi import CsError, Genome, ScreenCfg, Survivor, check, lib

SURVIVOR_DTYPE = np.dtype([("score", "<f4"), ("g", "<f4", (29,)), ("pa", "<u8"), ("pb", "<u8"),
                           ("tribe", "<u4"), ("stream", "<u4"), ("index", "<u8"), ("start", "<u8")])
assert SURVIVOR_DTYPE.itemsize == C.sizeof(Survivor)


class Model:
    """Flat quadratic surface (src/screen.zig:23-113)."""

    def __init__(self, raw: _abi.Model):
        self.raw = raw

    @classmethod
    def from_compact(cls, w: Sequence[float]) -> "Model":
        arr = np.ascontiguousarray(np.asarray(w, dtype=np.float32))
        m = _abi.Model()
        check(lib().cs_model_from_compact(arr.ctypes.data_as(C.POINTER(C.c_float)), int(arr.size), C.byref(m)))
        return cls(m)

    @classmethod
    def load(cls, path: str) -> "Model":
        m = _abi.Model()
        check(lib().cs_load_model(path.encode(), C.byref(m)))
        return cls(m)

    @property
    def w0(self) -> float:
        return float(self.raw.w0)

    @property
    def w_lin(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.raw.lin).copy()

    @property
    def w_quad(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.raw.quad).copy()


def load_model(path: str) -> Model:
    return Model.load(path)


@dataclass
class Cfg:
    """src/screen.zig:129-137 (threads fixes the RNG stream partition)."""
    lambda_: int
    keep: int
    seed: int
    threads: int = 4
    sigma: float = 0.12
    jump_p: float = 0.02
    imm_p: float = 0.01

    def to_c(self) -> ScreenCfg:
        return ScreenCfg(self.lambda_, self.keep, self.seed, self.threads,
                         self.sigma, self.jump_p, self.imm_p, 0.0)


def caller_cfg(screens: int, pop: int, seed: int, shard: int, gen: int, threads: int = 4, keep: int = 0,
               sigma_decay: float = 0.93) -> Cfg:
    """The screen parameters brain/evolve.py passes for generation `gen`
    (brain/evolve.py:368-371, 431-438; sigma_decay from brain/config.json:14):
    K = keep or pop - max(2, pop // 8) - max(2, pop // 8); seed = seed 1000003 +
    shard 131 + gen; sigma = "%.4f" % (0.12 sigma_decay^gen); jump "0.02";
    imm "0.01"; every value read as the Zig CLI reads its flags."""
    K = keep if keep > 0 else pop - max(2, pop // 8) - max(2, pop // 8)
    return Cfg(parse_u64(str(screens)), parse_u64(str(K)), parse_u64(str(seed * 1000003 + shard * 131 + gen)),
               parse_u64(str(threads)), parse_f32(f"{0.12 * (sigma_decay ** gen):.4f}"), parse_f32("0.02"),
               parse_f32("0.01"))


def parse_u64(tok: str) -> int:
    b = tok.encode()
    v = C.c_uint64()
    check(lib().cs_parse_u64(b, len(b), C.byref(v)))
    return v.value


def parse_f32(tok: str) -> float:
    b = tok.encode()
    v = C.c_float()
    check(lib().cs_parse_f32(b, len(b), C.byref(v)))
    return v.value


def parse_line(line: str) -> Optional[Genome]:
    """src/genome.zig:116-188: Genome, None for a skipped line; raises BadLine."""
    b = line.encode()
    g = Genome()
    rc = check(lib().cs_parse_pop_line(b, len(b), C.byref(g)))
    return g if rc == 1 else None


def load_elites(path: str) -> C.Array:
    """src/main.zig:201-213: every parseable line of the pop file."""
    n = C.c_int(0)
    rc = lib().cs_load_elites(path.encode(), None, 0, C.byref(n))
    if rc not in (0, -8):
        check(rc)
    arr = (Genome * max(1, n.value))()
    check(lib().cs_load_elites(path.encode(), arr, n.value, C.byref(n)))
    return (Genome * n.value).from_buffer(arr) if n.value else (Genome * 0)()


class Screen:
    """A device context running the funnel (one per host thread)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().cs_ctx_create(device, C.byref(h)))
        self._h = h
        self._children = weakref.WeakSet()  # MLP handles bound to this context

    def close(self) -> None:
        if self._h:
            for ch in list(self._children):
                ch.close()
            check(lib().cs_ctx_destroy(self._h))
            self._h = None

    def __enter__(self) -> "Screen":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, model: Model, elites, cfg: Cfg, noalloc: bool = False) -> Tuple[np.ndarray, int]:
        """Survivors (SURVIVOR_DTYPE, reference output order) and screens_count.
        noalloc: through cs_screen_noalloc (device buffers from the library's
        static scratch array instead of hipMalloc)."""
        ne = len(elites)
        if isinstance(elites, C.Array):
            earr = elites
        else:
            earr = (Genome * max(1, ne))(*elites)
        cap = max(1, min(int(cfg.keep), 1 << 26))
        out = np.zeros(cap, dtype=SURVIVOR_DTYPE)
        n_out = C.c_uint64(0)
        screens = C.c_uint64(0)
        c = cfg.to_c()
        fn = lib().cs_screen_noalloc if noalloc else lib().cs_screen
        check(fn(self._h, C.byref(model.raw), earr, ne, C.byref(c), out.ctypes.data_as(C.POINTER(Survivor)), cap,
                 C.byref(n_out), C.byref(screens)))
        return out[: n_out.value].copy(), screens.value

    def set_offset_mode(self, mode: int) -> None:
        """0 = automatic (speculative walk, verified), 1 = transfer maps."""
        check(lib().cs_set_offset_mode(self._h, int(mode)))

    def set_record_cap(self, cap: int) -> None:
        """Records per scoring-launch buffer (testing hook; 2048 <= cap <= 2^20)."""
        check(lib().cs_set_record_cap(self._h, int(cap)))

    def stats(self) -> dict:
        s = _abi.ScreenStats()
        check(lib().cs_get_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in _abi.ScreenStats._fields_}


def write_promoted(path: str, survivors: np.ndarray) -> None:
    arr = np.ascontiguousarray(survivors, dtype=SURVIVOR_DTYPE)
    check(lib().cs_write_promoted(path.encode(), arr.ctypes.data_as(C.POINTER(Survivor)), len(arr)))


def write_promoted_v6(path: str, survivors: np.ndarray) -> None:
    """The survivor lines with all 29 genes (34 tokens): the lines
    brain/evolve.py:445-454 accepts, unlike the reference's 29-token lines."""
    arr = np.ascontiguousarray(survivors, dtype=SURVIVOR_DTYPE)
    check(lib().cs_write_promoted_v6(path.encode(), arr.ctypes.data_as(C.POINTER(Survivor)), len(arr)))


def write_count(path: str, total: int) -> None:
    check(lib().cs_write_count(path.encode(), int(total)))


VALU_CLASSES = ("v_add_u32", "v_mul_lo_u32", "v_mul_hi_u32", "v_mad_u64_u32", "v_mul_f32", "v_lshl_add_u64",
                "v_alignbit_b32", "v_pk_fma_f16", "v_mfma_f32_16x16x32_bf16", "mfma+4pk_fma_f16",
                "v_mfma_f32_16x16x128_f8f6f4", "v_lshrrev_b64")


def valu_rates(device: int = 0, iters: int = 1 << 15) -> dict:
    """Measured issue cost of each instruction class at full occupancy: ns
    and s_memtime counts per instruction per SIMD (the last class: per MFMA,
    with four independent v_pk_fma_f16 issued beside each), the s_memtime
    rate those imply, and the SIMD count.  One untimed pass of the MFMA class
    first brings the device out of idle clocks."""
    ns_out, cyc_out = {}, {}
    simds = C.c_uint32(0)
    warm = C.c_double(0.0)
    check(lib().cs_valu_rate_probe(device, 8, iters, C.byref(warm), None, None))
    for op, name in enumerate(VALU_CLASSES):
        ns, cyc = C.c_double(0.0), C.c_double(0.0)
        check(lib().cs_valu_rate_probe(device, op, iters, C.byref(ns), C.byref(cyc), C.byref(simds)))
        ns_out[name] = ns.value / 8.0
        cyc_out[name] = cyc.value / 8.0
    ghz = sum(cyc_out.values()) / sum(ns_out.values())
    return {"ns_per_instr": ns_out, "memtime_per_instr": cyc_out, "memtime_ghz": ghz, "simds": simds.value}


def format_gene(v: float) -> str:
    buf = C.create_string_buffer(64)
    lib().cs_format_gene(C.c_float(v), buf, 64)
    return buf.value.decode()


V6_ENV = "CSFUNNEL_PROMOTED_FORMAT"


def screen_cmd(argv: List[str], device: int = 0) -> int:
    """`arena screen` (src/main.zig:156-218): flag parsing and defaults of the
    reference, model + elites loading, the screen itself, both output files and
    the stderr summary.  Returns the process exit code (1 on any error).

    Two additions, both off by default so the output files stay byte-identical
    to the reference's: `--v6-out FILE` also writes the survivors as 34-token
    lines with all 29 genes, and CSFUNNEL_PROMOTED_FORMAT=v6 in the
    environment writes `--out` itself that way -- the form brain/evolve.py
    (which passes no extra flags) keeps instead of dropping every line."""
    model_path = "state/surrogate.txt"
    v6_out = None
    elites_path = "out/elites.txt"
    out_path = "out/promoted.txt"
    count_out = "out/screens.txt"
    cfg = Cfg(lambda_=1_000_000, keep=192, seed=42)
    try:
        i = 0
        a = list(argv)
        while i < len(a):
            arg = a[i]
            if i + 1 >= len(a) and len(arg) > 0 and arg[0] == "-":
                raise CsError(-3, "flag without a value")
            if arg == "--model":
                i += 1
                model_path = a[i]
            elif arg == "--elites":
                i += 1
                elites_path = a[i]
            elif arg == "--out":
                i += 1
                out_path = a[i]
            elif arg == "--count-out":
                i += 1
                count_out = a[i]
            elif arg == "--lambda":
                i += 1
                cfg.lambda_ = parse_u64(a[i])
            elif arg == "--keep":
                i += 1
                cfg.keep = parse_u64(a[i])
            elif arg == "--seed":
                i += 1
                cfg.seed = parse_u64(a[i])
            elif arg == "--threads":
                i += 1
                cfg.threads = parse_u64(a[i])
            elif arg == "--sigma":
                i += 1
                cfg.sigma = parse_f32(a[i])
            elif arg == "--jump":
                i += 1
                cfg.jump_p = parse_f32(a[i])
            elif arg == "--imm":
                i += 1
                cfg.imm_p = parse_f32(a[i])
            elif arg == "--v6-out":
                i += 1
                v6_out = a[i]
            i += 1
        fmt = os.environ.get(V6_ENV, "")
        if fmt not in ("", "zig", "v6"):
            raise CsError(-3, f"{V6_ENV} must be 'zig' or 'v6', got {fmt!r}")
        model = load_model(model_path)
        elites = load_elites(elites_path)
        if len(elites) < 2:
            raise CsError(-2, "fewer than two elites")
        with Screen(device) as sc:
            surv, total = sc.run(model, elites, cfg)
        (write_promoted_v6 if fmt == "v6" else write_promoted)(out_path, surv)
        if v6_out is not None:
            write_promoted_v6(v6_out, surv)
        write_count(count_out, total)
    except CsError as e:
        sys.stderr.write(f"error: {e.name}\n{e}\n")
        return 1
    except OSError as e:
        sys.stderr.write(f"error: {e}\n")
        return 1
    sys.stderr.write(f"screen: {total} candidates scored, top {cfg.keep} promoted\n")
    return 0
