"""The deterministic synthetic corpus the codec labels are measured on.

Ten 320x192 YUV420p clips of 24 frames, generated from Python's Mersenne
Twister seeded with 0xC0DEC and drawn in a fixed clip order, exactly as
codec-arena's tools/gen_corpus.py defines them (clip recipes
tools/gen_corpus.py:29-232, chroma derivation :235-250, clip order and the
shared generator :253-268, file layout :283-298).  The files must hash to the
sha256 values of corpus/manifest.json; `generate` checks that and raises on
any mismatch, so a label is never computed on a corpus that differs from the
reference's.

Integer work and the plain IEEE double arithmetic of the recipes run
vectorised in numpy (same operations in the same order, each rounded once,
as in CPython); every transcendental call goes through CPython's `math`
module so the values are libm's, like the reference generator's.
"""
from __future__ import annotations

import hashlib
import math
import os
import random
from typing import Dict, List, Tuple

import numpy as np

W, H, FRAMES, FPS = 320, 192, 24, 24
SEED = 0xC0DEC
HOLDOUT = ("haze", "pulse")
# corpus/manifest.json of the reference snapshot
MANIFEST = {
    "drift": "c05a339b9b43a19b39bd47e45e03bb170f1d11c8a2908869f991725efa349606",
    "grid": "a2ec636530b0115ef0f651f8a921f9a2176619e62be78d143cbe692d3da08824",
    "storm": "a43678c2b04d219f1a83c59f45b4f8abde29484f727fbf4ee5b3e7e4bd08fe42",
    "quake": "5ebb9c8f3f8b1004d588297cbfab27ea5ee0a2aa7856f32c720eba38af7bf456",
    "grain": "b1ef714ec7a816519d9ae3f26c60816c5450e17552310e4d81d9b01e0f88ccb2",
    "text": "8881c8215efc77632f530293e1a7b3bae29ec0f7e9fb0fc1ca00443f93d1db2d",
    "zoom": "552f04dab8c0b25f0815b735a92af24b93436b3c5009d46f7b7cea13311ddcb2",
    "edge": "21c72590fd5d6cbc15634f78f64f556614fdfe54dfe4190fa704b3d7222ca14b",
    "haze": "ba7ffb364866cdf2047fa85208dda0c3bd7ef727b8cff7c8b14a8a3d1a2df251",
    "pulse": "7a2954bd9092b52c561547adb80d88c5ad2c591f36a358f238adbb8f38310614",
}
NAMES = tuple(MANIFEST)

_R = np.arange(H, dtype=np.int64)[:, None]
_C = np.arange(W, dtype=np.int64)[None, :]
_Rf = _R.astype(np.float64)
_Cf = _C.astype(np.float64)

_vsin = np.frompyfunc(math.sin, 1, 1)
_vexp = np.frompyfunc(math.exp, 1, 1)
_vsqrt = np.frompyfunc(math.sqrt, 1, 1)


def _lib(f, x):
    return f(x).astype(np.float64)


def _u8(v) -> np.ndarray:
    return np.clip(v, 0, 255).astype(np.uint8)


def _drift(rng) -> List[np.ndarray]:
    base = 60 + ((80 * _R) / H).astype(np.int64)          # int(80 * r / H)
    out = []
    for t in range(FRAMES):
        bx = int(W * (0.5 + 0.35 * math.sin(2 * math.pi * t / FRAMES)))
        by = int(H * (0.5 + 0.35 * math.cos(2 * math.pi * t / FRAMES)))
        d2 = (_C - bx) ** 2 + (_R - by) ** 2
        v = np.where(d2 < 900, np.clip(base + 70 - d2 // 30, 0, 255), base + 0 * d2)
        out.append(_u8(v))
    return out


def _grid(rng) -> List[np.ndarray]:
    out = []
    for t in range(FRAMES):
        shift = (t * 4) % 32
        pulse = 90 + int(60 * math.sin(2 * math.pi * t / FRAMES))
        cell = ((_C + shift) // 16 + _R // 16) % 2
        out.append(_u8(np.where(cell == 0, pulse, 255 - pulse)))
    return out


def _storm(rng) -> List[np.ndarray]:
    base = np.array([rng.random() for _ in range(64 * 64)], dtype=np.float64)
    out = []
    for t in range(FRAMES):
        ox, oy = t * 0.7, t * 0.35
        fx = np.mod(_Cf * 0.2 + ox, 64)
        fy = np.mod(_Rf * 0.2 + oy, 64)
        x0, y0 = np.trunc(fx).astype(np.int64), np.trunc(fy).astype(np.int64)
        x1, y1 = (x0 + 1) % 64, (y0 + 1) % 64
        tx, ty = fx - x0, fy - y0
        a, b = base[y0 * 64 + x0], base[y0 * 64 + x1]
        cc, d = base[y1 * 64 + x0], base[y1 * 64 + x1]
        v = (a * (1 - tx) + b * tx) * (1 - ty) + (cc * (1 - tx) + d * tx) * ty
        out.append(np.trunc(v * 255).astype(np.int64).astype(np.uint8))
    return out


def _quake(rng) -> List[np.ndarray]:
    out = []
    ax, ay, avx, avy = 30.0, 40.0, 9.0, 6.0
    bx, by, bvx, bvy = 230.0, 100.0, -8.0, 5.0
    for t in range(FRAMES):
        shake = ((t * 37) % 9) - 4
        sx = np.array([int(50 * math.sin((c + shake) * 0.05)) for c in range(W)], dtype=np.int64)[None, :]
        cy = np.array([int(20 * math.cos((r + shake) * 0.07)) for r in range(H)], dtype=np.int64)[:, None]
        v = 90 + sx + cy
        da = (_C - int(ax)) ** 2 + (_R - int(ay)) ** 2
        db = (_C - int(bx)) ** 2 + (_R - int(by)) ** 2
        v = np.where(da < 1600, 235 - da // 40, v)
        v = np.where(db < 1200, 20 + db // 30, v)
        out.append(_u8(v))
        ax += avx
        ay += avy
        if ax < 20 or ax > W - 20:
            avx = -avx
        if ay < 20 or ay > H - 20:
            avy = -avy
        bx += bvx
        by += bvy
        if bx < 20 or bx > W - 20:
            bvx = -bvx
        if by < 20 or by > H - 20:
            bvy = -bvy
    return out


def _grain(rng) -> List[np.ndarray]:
    out = []
    for t in range(FRAMES):
        lo, hi = 110 + (t % 3) * 12, 150 + (t % 5) * 8
        n = hi - lo
        out.append(np.array([lo + rng.randrange(n) for _ in range(W * H)], dtype=np.uint8).reshape(H, W))
    return out


def _text(rng) -> List[np.ndarray]:
    glyphs = np.array([[rng.random() < 0.55 for _ in range(20)] for _ in range(8)], dtype=bool)
    stroke_r = np.isin(_R % 24, (3, 4, 9, 10, 15, 16))
    backdrop = np.where((_R // 24) % 2 == 0, 128, 60)
    out = []
    for t in range(FRAMES):
        shift = (t * 3) % 40
        band = _R // 24
        col = (_C + shift) // 20
        cell_c = (_C + shift) % 20
        on = glyphs[(band * 3 + col) % 8, col % 20]
        stroke = on & stroke_r & np.isin(cell_c, (2, 3, 6, 7, 10, 11, 14))
        v = np.where(stroke, 235, np.where(on, 45, backdrop))
        out.append(_u8(v))
    return out


def _zoom(rng) -> List[np.ndarray]:
    base = np.array([[rng.random() for _ in range(80)] for _ in range(48)], dtype=np.float64)
    out = []
    cx0, cy0 = W / 2.0, H / 2.0
    for t in range(FRAMES):
        z = 1.0 + 0.35 * (t / FRAMES)
        fx = (_Cf - cx0) / z + cx0
        fy = (_Rf - cy0) / z + cy0
        fx = np.mod(fx * 0.25, 79.0)
        fy = np.mod(fy * 0.25, 47.0)
        fx, fy = np.broadcast_arrays(fx, fy)
        x0, y0 = np.trunc(fx).astype(np.int64), np.trunc(fy).astype(np.int64)
        tx, ty = fx - x0, fy - y0
        x1, y1 = (x0 + 1) % 80, (y0 + 1) % 48
        a = base[y0, x0] * (1 - tx) + base[y0, x1] * tx
        b = base[y1, x0] * (1 - tx) + base[y1, x1] * tx
        v = a * (1 - ty) + b * ty
        checker = ((x0 // 4 + y0 // 4) % 2) * 0.4
        out.append(_u8(np.trunc(255 * (0.3 + 0.5 * v + checker * (v - 0.5))).astype(np.int64)))
    return out


def _edge(rng) -> List[np.ndarray]:
    out = []
    for t in range(FRAMES):
        ang = 2 * math.pi * t / FRAMES
        proj = (_Cf - W / 2) * math.cos(ang) + (_Rf - H / 2) * math.sin(ang)
        s = _lib(_vsin, proj * 0.22)
        v = 128 + 90 * s
        v = np.where(np.abs(s) > 0.93, np.where(v > 128, 255.0, 0.0), v)
        out.append(_u8(np.trunc(v).astype(np.int64)))
    return out


def _haze(rng) -> List[np.ndarray]:
    layers = [[(rng.random(), rng.random()) for _ in range(2)] for _ in range(3)]
    out = []
    for t in range(FRAMES):
        v = 60 + (70 * _R) / H + 0.0 * _Cf
        for li in range(3):
            p = layers[li]
            sgx = 1 if p[0][0] > 0.5 else -1
            sgy = 1 if p[0][1] > 0.5 else -1
            fx = np.mod(_Cf * 0.01 + t * (0.05 + 0.03 * li) * sgx, 2)
            fy = np.mod(_Rf * 0.01 + t * 0.02 * sgy, 2)
            wgt = 12 * (li + 1) / 3
            sx = _lib(_vsin, 6.28 * fx + p[1][0] * 6.28)
            sy = _lib(_vsin, 6.28 * fy + p[1][1] * 6.28)
            v = v + wgt * (0.5 + 0.5 * sx) * (0.5 + 0.5 * sy)
        out.append(_u8(np.trunc(v).astype(np.int64)))
    return out


def _pulse(rng) -> List[np.ndarray]:
    d = _lib(_vsqrt, (_Cf - W / 2) ** 2 + (_Rf - H / 2) ** 2)
    ed = _lib(_vexp, -d / 200)
    out = []
    for t in range(FRAMES):
        phase = 2 * math.pi * t / FRAMES
        v = 128 + 80 * _lib(_vsin, d * 0.3 - 3 * phase) * ed
        out.append(_u8(np.trunc(v).astype(np.int64)))
    return out


_RECIPES = (("drift", _drift), ("grid", _grid), ("storm", _storm), ("quake", _quake), ("grain", _grain),
            ("text", _text), ("zoom", _zoom), ("edge", _edge), ("haze", _haze), ("pulse", _pulse))


def _chroma(y: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    yi = y.astype(np.int64)
    m = (yi[0::2, 0::2] + yi[0::2, 1::2] + yi[1::2, 0::2] + yi[1::2, 1::2]) // 4
    return _u8(128 + (m - 128) // 4), _u8(128 - (m - 128) // 4)


def clip_bytes(frames: List[np.ndarray]) -> bytes:
    """Global-planar layout: all Y frames, then all U, then all V."""
    us, vs = zip(*(_chroma(f) for f in frames))
    return b"".join(f.tobytes() for f in frames) + b"".join(u.tobytes() for u in us) + \
        b"".join(v.tobytes() for v in vs)


def generate(out_dir: str) -> Dict[str, str]:
    """Write <name>.yuv for all ten clips plus corpus.txt into out_dir and
    return their sha256; raises if any differs from the reference manifest."""
    os.makedirs(out_dir, exist_ok=True)
    rng = random.Random(SEED)
    hashes = {}
    for name, fn in _RECIPES:
        data = clip_bytes(fn(rng))
        hashes[name] = hashlib.sha256(data).hexdigest()
        if hashes[name] != MANIFEST[name]:
            raise RuntimeError(f"corpus clip {name}: sha256 {hashes[name]} != manifest {MANIFEST[name]}")
        with open(os.path.join(out_dir, name + ".yuv.tmp"), "wb") as f:
            f.write(data)
        os.replace(os.path.join(out_dir, name + ".yuv.tmp"), os.path.join(out_dir, name + ".yuv"))
    with open(os.path.join(out_dir, "corpus.txt"), "w") as f:
        f.write("".join(f"{n} {W} {H} {FRAMES} {FPS}\n" for n in NAMES))
    return hashes


def default_dir() -> str:
    return os.environ.get("CSFUNNEL_CORPUS") or os.path.join(
        os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "corpus")


def ensure(out_dir: str = None) -> str:
    """Return a directory holding the verified corpus, generating it once."""
    out_dir = out_dir or default_dir()
    ok = all(os.path.exists(os.path.join(out_dir, n + ".yuv")) for n in NAMES)
    if ok:
        for n in NAMES:
            with open(os.path.join(out_dir, n + ".yuv"), "rb") as f:
                if hashlib.sha256(f.read()).hexdigest() != MANIFEST[n]:
                    ok = False
                    break
    if not ok:
        generate(out_dir)
    return out_dir


def load(out_dir: str = None, names=NAMES) -> Dict[str, bytes]:
    d = ensure(out_dir)
    res = {}
    for n in names:
        with open(os.path.join(d, n + ".yuv"), "rb") as f:
            res[n] = f.read()
    return res
