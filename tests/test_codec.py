"""The codec label source (CPU): corpus, encoder/decoder, metrics.

Pinned by the reference's own data: the corpus against corpus/manifest.json;
all 22 hall-of-fame records against their published PSNR, bpp, dec_mops,
chroma PSNR, MS-SSIM and MS-SSIM dB (halloffame/*.json, measured by the
reference on the five legacy clips of tools/gen_corpus.py); the one
leaderboard genome against its row on the ten-clip corpus.  Beside them: the
closed-loop invariant of src/codec.zig:2833-2970, the unit-test known answers
of src/codec.zig:2973-3124 (sub-pel geometry, intra prediction, deblocking),
the range coder round trip of :3001-3094, the MS-SSIM and PSNR properties of
src/msssim.zig:186-226 and src/yuv.zig:82-85, and an FP64 MS-SSIM written
from Wang et al. 2003.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

LEGACY5 = ("drift", "grid", "storm", "quake", "grain")
PRINTED = {"psnr": 2, "bpp": 4, "dec_mops": 2, "psnr_u": 2, "psnr_v": 2, "msssim": 5, "msssim_db": 4}


@pytest.fixture(scope="module")
def cc():
    from csfunnel import codec
    codec.lib()
    return codec


@pytest.fixture(scope="module")
def clips(tmp_path_factory):
    from csfunnel import corpus
    d = str(tmp_path_factory.mktemp("corpus"))
    corpus.generate(d)
    return d


def test_corpus_matches_manifest(clips):
    from csfunnel import corpus
    for n in corpus.NAMES:
        with open(os.path.join(clips, n + ".yuv"), "rb") as f:
            data = f.read()
        assert len(data) == corpus.FRAMES * corpus.W * corpus.H * 3 // 2
        assert hashlib.sha256(data).hexdigest() == corpus.MANIFEST[n]
    assert open(os.path.join(clips, "corpus.txt")).read().splitlines()[0] == "drift 320 192 24 24"


def _kat_genome(cc, r):
    # the oracle reads genomes from pop lines written with "%.4f" (brain/evolve.py:104-106)
    return cc.genome([float("%.4f" % x) for x in r["g"]], r["id"])


def test_halloffame_records_reproduce(cc, clips):
    kats = json.load(open(os.path.join(GOLDEN, "halloffame_kats.json")))
    recs = kats["records"] + [kats["champion"]]
    ev = cc.Evaluator(names=LEGACY5, corpus_dir=clips, msssim=True)
    gs = [_kat_genome(cc, r) for r in recs]
    assert all(cc.supported(g) for g in gs)
    ms = ev.evaluate(gs)
    checked = 0
    for r, m in zip(recs, ms):
        ref = r["metrics"]
        for k, d in PRINTED.items():
            if k in ref:
                assert round(m[k], d) == pytest.approx(ref[k], abs=0.51 * 10 ** -d), (r["id"], k, m[k], ref[k])
                checked += 1
    assert checked >= 110


def test_leaderboard_bot_reproduces_on_the_full_corpus(cc, clips):
    """Bot 13 of pop/seed.txt read through the reference parser (its 34-token
    line loses genes 16-21, so the entropy gene is 0): the byte-varint stream
    with its drifting decode, on all ten clips, gives the LEADERBOARD.md row."""
    from csfunnel import screen as scr
    kat = json.load(open(os.path.join(GOLDEN, "halloffame_kats.json")))["leaderboard"]
    g = scr.parse_line(kat["pop_line"])
    assert list(g.g)[16:22] == [0.0] * 6 and cc.supported(g)
    m = cc.Evaluator(corpus_dir=clips, msssim=True).metrics(g)
    for k, d in (("psnr", 2), ("bpp", 4), ("msssim", 4)):
        assert round(m[k], d) == pytest.approx(kat["metrics"][k], abs=0.51 * 10 ** -d), (k, m[k])


def _synth_clip(w, h, frames):
    """synthClip (src/codec.zig:2775-2809): a gradient drifting 2 px right
    and 1 px down per frame, a 16x16 box of 220 and a 6x6 box of 30; flat
    chroma 100 / 150.  Planar Y frames, then U, then V."""
    r = np.arange(h)[:, None]
    c = np.arange(w)[None, :]
    ys = []
    for fi in range(frames):
        dx, dy = (c + fi * 2) % w, (r + fi) % h
        px = ((dx * 255) // w + (dy * 255) // h) // 2
        px = np.where((dx >= 6) & (dx < 22) & (dy >= 6) & (dy < 22), 220, px)
        px = np.where((dx >= 24) & (dx < 30) & (dy >= 4) & (dy < 10), 30, px)
        ys.append(np.clip(px, 0, 255).astype(np.uint8))
    n = frames * w * h // 4
    return np.stack(ys).tobytes() + bytes([100]) * n + bytes([150]) * n


def _v8(intra, deb, entropy, cfl, cdef, qpel, sb16):
    """genomeWithV8 (src/codec.zig:2811-2831, 3126-3130): every gene 0.5,
    baseq 0.18, the v0.4-v0.8 tool genes as given."""
    g = [0.5] * 29
    g[3] = 0.18
    g[16:23] = [intra, deb, entropy, cfl, cdef, qpel, sb16]
    return g


def test_superblock_gene_at_or_below_half_is_the_legacy_stream(cc):
    """src/codec.zig:3207-3225: g[22] <= 0.5 never writes a superblock flag,
    so 0, 0.4 and 0.5 code the same stream; above 0.5 the flags are there."""
    data = _synth_clip(64, 64, 4)
    ref = None
    for v in (0.0, 0.4, 0.5):
        st = cc.eval_clip(data, cc.genome(_v8(0.9, 0.6, 0.9, 0.9, 0.9, 0.9, v), 11), w=64, h=64, frames=4,
                          flags=cc.FAST_SAD | cc.CLOSED_LOOP)
        got = (st.bytes, st.psnr_y, st.psnr_u, st.dct_calls, st.idct_calls, tuple(st.sb_unified))
        assert got[-1] == (0, 0, 0)
        assert ref is None or got == ref
        ref = got
    narrow = _synth_clip(16, 64, 4)
    off = cc.eval_clip(narrow, cc.genome(_v8(0.9, 0.6, 0.9, 0.9, 0.9, 0.9, 0.5), 11), w=16, h=64, frames=4)
    on = cc.eval_clip(narrow, cc.genome(_v8(0.9, 0.6, 0.9, 0.9, 0.9, 0.9, 0.51), 11), w=16, h=64, frames=4)
    assert on.bytes > off.bytes and on.closed_loop == off.closed_loop == 1


def _sb_genomes(n, seed=5):
    rng = np.random.default_rng(seed)
    for t in range(n):
        g = list(np.round(rng.random(29), 4))
        g[18] = 0.9
        g[22] = float(np.round(0.51 + 0.49 * rng.random(), 4))
        yield t, g


def test_superblock_closed_loop_with_unified_superblocks(cc):
    """THE v0.8 invariant (src/codec.zig:3174-3203): with the superblock
    layer on, the decoder output equals the encoder recon byte for byte.
    Planes two blocks wide (16-pixel luma, 32-pixel clips' chroma) are the
    ones where a unified superblock can win and be committed; over 200
    random range-coded genomes all three unified subtypes -- intra16, skip16,
    inter16 -- are committed and decoded, and every clip either closes the
    loop or meets one of the two undefined events."""
    clips = {(w, h): _synth_clip(w, h, 6) for w, h in ((16, 64), (16, 32), (32, 32))}
    seen = np.zeros(3, dtype=np.int64)
    ok = unsupported = 0
    for t, g in _sb_genomes(200):
        gn = cc.genome(g, t)
        for (w, h), data in clips.items():
            try:
                st = cc.eval_clip(data, gn, w=w, h=h, frames=6, flags=cc.FAST_SAD | cc.CLOSED_LOOP)
            except cc.CodecError as e:
                assert e.code == -20, (t, w, h, str(e))
                unsupported += 1
                continue
            assert st.closed_loop == 1
            seen += np.array(st.sb_unified)
            ok += 1
    assert (seen > 0).all(), seen
    assert ok > 400 and unsupported > 50, (ok, unsupported)


def test_superblock_rollback_on_the_reference_test_clip(cc):
    """The four genomes of the reference's superblock closed-loop test on its
    64x64 clip (src/codec.zig:3178-3191).  In each, the unified trial of a row
    pair's last superblock wins in a P frame of the 8-block-wide luma plane,
    so the reference's rollback drops six coded blocks from the stream: the
    label source refuses the clip instead of decoding it."""
    data = _synth_clip(64, 64, 6)
    for i, g in enumerate(((0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.9), (0.9, 0.6, 0.9, 0.9, 0.9, 0.9, 0.6),
                           (0.9, 0.0, 0.0, 0.9, 0.0, 0.0, 1.0), (0.0, 0.6, 0.9, 0.0, 0.9, 0.9, 0.75))):
        with pytest.raises(cc.CodecError) as e:
            cc.eval_clip(data, cc.genome(_v8(*g), i + 1), w=64, h=64, frames=6)
        assert e.value.code == -20
        # the first frame is an I frame where legacy keeps every superblock
        st = cc.eval_clip(data, cc.genome(_v8(*g), i + 1), w=64, h=64, frames=1)
        assert tuple(st.sb_unified) == (0, 0, 0)


def test_seed_population_superblock_genomes(cc, clips):
    """pop/seed.txt holds 61 genomes with g[22] > 0.5.  Read through the
    reference's pop-line parser (its 34-token lines lose genes 16-21, so the
    byte-varint stream codes them), none of the 61 evaluates on any of the
    ten 320x192 clips: in each, the last superblock of some row pair wins
    its trial against a rate that counts every block coded since its
    snapshot (38 more on the 40-block luma plane), and the rollback is
    undefined."""
    from csfunnel import corpus, screen as scr
    data = corpus.load(clips)
    gs = [scr.parse_line(ln) for ln in open(os.path.join(GOLDEN, "seed_pop.txt")).read().splitlines()]
    sb = [g for g in gs if g.g[22] > 0.5]
    assert len(gs) == 96 and len(sb) == 61
    evaluated = 0
    for g in sb:
        for n in corpus.NAMES:
            try:
                cc.eval_clip(data[n], g)
                evaluated += 1
            except cc.CodecError as e:
                assert e.code == -20
    assert evaluated == 0
    # the other 35 seed genomes evaluate
    for g in gs:
        if g.g[22] <= 0.5:
            cc.eval_clip(data["drift"], g, frames=2)


def test_closed_loop_holds_on_the_range_coder_stream(cc, clips):
    """Decoder output == encoder recon, byte for byte, with the adaptive range
    coder on and every other tool in combination: the intra mode tiers,
    chroma-from-luma, CDEF, quarter-pel or half-pel MC, deblocking, long
    GOPs, intra refresh and adaptive q (src/codec.zig:2833-2970)."""
    from csfunnel import corpus
    data = corpus.load(clips)
    rng = np.random.default_rng(11)
    for i in range(10):
        g = list(np.round(rng.random(16), 4)) + [0.0] * 7 + list(np.round(rng.random(6), 4))
        g[12] = 0.9 if i % 2 == 0 else 0.1
        g[16] = (0.1, 0.4, 0.7, 0.9, 0.95)[i % 5]
        g[17], g[18] = float(rng.random()), 0.8
        g[19] = 0.9 if i % 2 else 0.2
        g[20] = 0.9 if i % 3 == 0 else 0.2
        g[21] = 0.9 if i % 4 < 2 else 0.2
        gn = cc.genome(g, i)
        for n in ("quake", "edge"):
            st = cc.eval_clip(data[n], gn, flags=cc.FAST_SAD | cc.CLOSED_LOOP)
            assert st.closed_loop == 1 and st.idct_calls == 2 * st.dct_calls


def test_interp_q_geometry(cc):
    """src/codec.zig:2973-2999: every sub-pel position of a flat plane is the
    flat value; the +2 (half) position of a linear ramp is the ramp midpoint;
    the half-plane path equals the direct filters away from the border."""
    import ctypes as C
    L = cc.lib()
    out = C.create_string_buffer(64)
    flat = bytes([100]) * 1600
    for qx in range(-2, 3):
        for qy in range(-2, 3):
            for direct in (0, 1):
                assert L.csc_interp_q(flat, 40, 40, 8, 8, qx, qy, direct, out) == 0
                assert set(out.raw) == {100}
    ramp = bytes([50 + c * 2 for r in range(40) for c in range(40)])
    for direct in (0, 1):
        assert L.csc_interp_q(ramp, 40, 40, 8, 8, 2, 0, direct, out) == 0
        assert list(out.raw) == [50 + (8 + c) * 2 + 1 for r in range(8) for c in range(8)]
    noise = np.random.default_rng(2).integers(0, 256, 48 * 48).astype(np.uint8).tobytes()
    o2 = C.create_string_buffer(64)
    for qx in range(-2, 3):
        for qy in range(-2, 3):
            L.csc_interp_q(noise, 48, 48, 16, 20, qx, qy, 0, out)
            L.csc_interp_q(noise, 48, 48, 16, 20, qx, qy, 1, o2)
            assert out.raw == o2.raw, (qx, qy)


def test_intra_predict_known_answers(cc):
    """src/codec.zig:3095-3109 (vertical copies the top row, DC of a flat
    plane is flat), plus horizontal, the NONE mode's mid-gray, the 128 fill
    at the frame corner and chroma-from-luma on flat luma."""
    import ctypes as C
    L = cc.lib()
    out = C.create_string_buffer(64)
    rec = bytes([(r * 7 + c * 3) % 256 for r in range(32) for c in range(32)])
    assert L.csc_intra_predict(3, rec, 32, 32, 8, 8, None, 0, out) == 0
    assert list(out.raw) == [rec[7 * 32 + 8 + c] for r in range(8) for c in range(8)]
    assert L.csc_intra_predict(4, rec, 32, 32, 8, 8, None, 0, out) == 0
    assert list(out.raw) == [rec[(8 + r) * 32 + 7] for r in range(8) for c in range(8)]
    assert L.csc_intra_predict(1, bytes([90]) * 1024, 32, 32, 8, 8, None, 0, out) == 0
    assert set(out.raw) == {90}
    assert L.csc_intra_predict(0, rec, 32, 32, 8, 8, None, 0, out) == 0 and set(out.raw) == {128}
    assert L.csc_intra_predict(3, rec, 32, 32, 0, 0, None, 0, out) == 0 and set(out.raw) == {128}
    luma = bytes([200]) * (64 * 64)
    for mode in range(10, 14):
        assert L.csc_intra_predict(mode, bytes([77]) * 1024, 32, 32, 8, 8, luma, 64, out) == 0
        assert set(out.raw) == {77}
    assert L.csc_intra_predict(10, rec, 32, 32, 8, 8, None, 0, out) < 0


def test_deblock_pair_known_answers(cc):
    """src/codec.zig:3111-3124: the midpoint is never crossed, flat stays flat."""
    import ctypes as C
    L = cc.lib()
    p0, q0 = C.c_uint8(100), C.c_uint8(110)
    L.csc_deblock_pair(C.byref(p0), C.byref(q0), 100, 110, 50.0, 25.0)
    assert 100 < p0.value <= 105 and 105 <= q0.value < 110
    a, b = C.c_uint8(77), C.c_uint8(77)
    L.csc_deblock_pair(C.byref(a), C.byref(b), 77, 77, 50.0, 25.0)
    assert a.value == b.value == 77


def test_varint_stream_reproduces_the_reference_decode(cc, clips):
    """The byte-varint decoder reads a negative x back as x + 1
    (src/codec.zig:215-220), so that stream's decoder drifts from the encoder."""
    L = cc.lib()
    for v in (0, 1, 2, 63, 64, 1 << 40):
        assert L.csc_svarint_roundtrip(v) == v
    for v in (-1, -2, -3, -64, -65, -(1 << 40)):
        assert L.csc_svarint_roundtrip(v) == v + 1
    from csfunnel import corpus
    data = corpus.load(clips, ("drift",))["drift"]
    g = [0.3] * 16
    st = cc.eval_clip(data, cc.genome(g))
    assert st.closed_loop == 0
    with pytest.raises(cc.CodecError):
        cc.eval_clip(data, cc.genome(g), flags=cc.FAST_SAD | cc.CLOSED_LOOP)
    g_on = cc.genome(g[:16] + [0.0, 0.0, 0.9, 0.0, 0.0, 0.0, 0.0])
    on = cc.eval_clip(data, g_on, flags=cc.FAST_SAD | cc.CLOSED_LOOP)
    assert on.closed_loop == 1 and on.bytes < st.bytes


def test_range_coder_roundtrip(cc):
    import ctypes as C
    L = cc.lib()
    rng = np.random.default_rng(5)
    for n, p1 in ((0, 0.5), (1, 0.5), (5000, 0.5), (20000, 0.02), (20000, 0.98)):
        bits = (rng.random(n) < p1).astype(np.uint8)
        ctx = rng.integers(0, 64, n).astype(np.uint8)
        buf = C.create_string_buffer(n + 64)
        back = C.create_string_buffer(max(1, n))
        m = L.csc_rc_roundtrip(bits.tobytes(), ctx.tobytes(), n, buf, n + 64, back)
        assert m >= 5
        assert np.frombuffer(back.raw[:n], dtype=np.uint8).tolist() == bits.tolist()
        if n == 0:
            assert m == 5
        if n == 20000:
            # adaptive probabilities: skewed streams cost well under a bit each
            h = -(p1 * np.log2(p1) + (1 - p1) * np.log2(1 - p1))
            assert m * 8 < n * (h + 0.1)


def _msssim_f64(a, b):
    """MS-SSIM (Wang, Simoncelli, Bovik 2003) in FP64: valid-mode 11-tap
    gaussian sigma 1.5, 2x2 mean pooling, weights (.0448 .2856 .3001 .2363
    .1333), C1 = (0.01 L)^2, C2 = (0.03 L)^2."""
    x = np.arange(11) - 5.0
    k = np.exp(-x * x / 4.5)
    k /= k.sum()

    def filt(im):
        h, w = im.shape
        t = sum(k[i] * im[:, i:i + w - 10] for i in range(11))
        return sum(k[i] * t[i:i + h - 10, :] for i in range(11))

    W = [0.0448, 0.2856, 0.3001, 0.2363, 0.1333]
    X, Y = a.astype(np.float64), b.astype(np.float64)
    out = 0.0
    for s in range(5):
        mx, my = filt(X), filt(Y)
        vx, vy, vxy = filt(X * X) - mx * mx, filt(Y * Y) - my * my, filt(X * Y) - mx * my
        cs = ((2 * vxy + 58.5225) / (vx + vy + 58.5225)).mean()
        if s == 4:
            out += W[s] * np.log(((2 * mx * my + 6.5025) / (mx * mx + my * my + 6.5025)).mean())
        out += W[s] * np.log(cs)
        if s < 4:
            h, w = X.shape
            X = (X[0:h // 2 * 2:2, 0:w // 2 * 2:2] + X[0:h // 2 * 2:2, 1:w // 2 * 2:2] + X[1:h // 2 * 2:2, 0:w // 2 * 2:2]
                 + X[1:h // 2 * 2:2, 1:w // 2 * 2:2]) * 0.25
            Y = (Y[0:h // 2 * 2:2, 0:w // 2 * 2:2] + Y[0:h // 2 * 2:2, 1:w // 2 * 2:2] + Y[1:h // 2 * 2:2, 0:w // 2 * 2:2]
                 + Y[1:h // 2 * 2:2, 1:w // 2 * 2:2]) * 0.25
    return float(np.exp(out))


def _ms(cc, a, b):
    import ctypes as C
    h, w = a.shape
    v = C.c_float()
    assert cc.lib().csc_plane_msssim(a.astype(np.uint8).tobytes(), b.astype(np.uint8).tobytes(), w, h, C.byref(v)) == 0
    return v.value


def test_msssim_properties_and_fp64(cc):
    i = np.arange(64 * 64)
    buf = ((i * 7) % 256).reshape(64, 64)
    assert _ms(cc, buf, buf) > 0.9999
    base = ((i * 3 + (i // 64) * 5) % 256).reshape(64, 64)
    rng = np.random.default_rng(0xC0DEC)
    mild = np.minimum(255, np.maximum(0, base + rng.integers(-4, 5, base.shape)))
    heavy = np.minimum(255, np.maximum(0, base + rng.integers(-40, 41, base.shape)))
    m, hv = _ms(cc, base, mild), _ms(cc, base, heavy)
    assert m > hv and 0.05 < hv < 0.999
    a = ((np.arange(32 * 32) * 11) % 200 + 20).reshape(32, 32)
    v = _ms(cc, a, a + 10)
    assert 0.90 < v < 1.0
    for (h, w) in ((192, 320), (176, 176)):
        x = rng.integers(0, 256, (h, w))
        y = np.clip(x + rng.integers(-20, 21, (h, w)), 0, 255)
        assert _ms(cc, x, y) == pytest.approx(_msssim_f64(x, y), abs=2e-5)


def test_psnr_plane(cc):
    a = np.arange(1000, dtype=np.uint8)
    assert cc.lib().csc_psnr_plane(a.tobytes(), a.tobytes(), 1000) == 99.0
    b = (a.astype(np.int64) + np.where(np.arange(1000) % 2 == 0, 3, -3)).clip(0, 255).astype(np.uint8)
    mse = ((a.astype(np.float64) - b) ** 2).mean()
    assert cc.lib().csc_psnr_plane(a.tobytes(), b.tobytes(), 1000) == pytest.approx(
        10 * np.log10(255.0 ** 2 / mse), abs=1e-5)


def test_codec_library_exports_every_declared_symbol(cc):
    import re
    import ctypes as C
    txt = open(os.path.join(os.path.dirname(GOLDEN), "..", "include", "cscodec.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = sorted(set(re.findall(r"\b(csc_[a-z0-9_]+)\s*\(", txt)))
    assert len(names) == 10
    L = cc.lib()
    for n in names:
        assert hasattr(L, n) and n in cc.SIGNATURES, n
    # ... + closed_loop, enc_ms, sb_unified[3], then 4 bytes of tail padding to 8-byte alignment
    assert C.sizeof(cc.ClipStats) == 8 + 6 * 4 + 8 + 8 + 4 + 4 + 3 * 4 + 4


# ------------------------------------------------------------ arena eval CLI

def test_zig_decimal_matches_the_oracle_printer(oracle):
    """{d:.4} of f32 against the oracle's printer (src/screen.zig:330 format),
    including values below 1e-4 and round-half-up carries."""
    from csfunnel.evalcmd import zig_decimal
    rng = np.random.default_rng(11)
    vals = np.concatenate([rng.random(2000), rng.random(500) * 1e-3, rng.random(300) * 1e4,
                           -rng.random(200), [0.0, 0.99996, 0.99995, 9.99995, 5e-05, 4e-05, 1e-45]])
    for v in vals.astype(np.float32):
        assert zig_decimal(float(v), 4) == oracle.fmt(float(v)), repr(v)


@pytest.mark.parametrize("v,prec,bits,s", [(31.675, 2, 32, "31.68"), (0.18155, 4, 64, "0.1816"),
                                           (1784.25, 1, 32, "1784.3"), (0.0, 2, 32, "0.00"),
                                           (0.828749, 5, 32, "0.82875"), (12.5, 6, 32, "12.500000"),
                                           (0.99999, 2, 32, "1.00"), (-3.14159, 2, 64, "-3.14")])
def test_zig_decimal_known(v, prec, bits, s):
    from csfunnel.evalcmd import zig_decimal
    assert zig_decimal(v, prec, bits) == s


def _corpus_subset(clips, d, names):
    os.makedirs(d, exist_ok=True)
    for n in names:
        dst = os.path.join(d, n + ".yuv")
        if not os.path.exists(dst):
            os.symlink(os.path.join(clips, n + ".yuv"), dst)
    with open(os.path.join(d, "corpus.txt"), "w") as f:
        f.write("# name w h frames fps\n" + "".join(f"{n} 320 192 24 24\n" for n in names))
    return d


def test_arena_eval_writes_the_halloffame_rows(cc, clips, tmp_path, capfd):
    """`arena eval` on the five legacy clips over a population file of all 22
    hall-of-fame genomes (33-token v5 lines, "%.4f" genes) prints their
    published PSNR, bpp, dec_mops, chroma PSNR, MS-SSIM and MS-SSIM dB
    columns digit for digit; a 35-token line and a superblock genome are
    counted bad, a 3-token line is skipped (src/main.zig:279-355)."""
    from csfunnel.arena import main
    kats = json.load(open(os.path.join(GOLDEN, "halloffame_kats.json")))
    recs = kats["records"] + [kats["champion"]]
    corp = _corpus_subset(clips, str(tmp_path / "corpus"), LEGACY5)
    lines = [f"{r['id']} 0 0 0 0 " + " ".join("%.4f" % x for x in r["g"]) for r in recs]
    sb16 = "99 0 0 0 0 " + " ".join(["0.5000"] * 22) + " 0.9000 " + " ".join(["0.5000"] * 6)
    pop = tmp_path / "pop.txt"
    pop.write_text("\n".join(lines[:3] + ["1 2 3", " ".join(["1"] * 35), sb16] + lines[3:]) + "\n")
    out = tmp_path / "res.csv"
    assert main(["eval", "--corpus", corp, "--pop", str(pop), "--out", str(out)]) == 0
    err = capfd.readouterr().err
    assert f"eval: {len(recs)} ok, 2 bad, 5 clips" in err and "[evalClip error: error.Unsupported clip=drift]" in err
    rows = [ln.split(",") for ln in out.read_text().splitlines()]
    assert [int(t[0]) for t in rows] == [r["id"] for r in recs]
    cols = {"psnr": 1, "bpp": 2, "dec_mops": 4, "psnr_u": 5, "psnr_v": 6, "msssim": 7, "msssim_db": 8}
    checked = 0
    for r, t in zip(recs, rows):
        assert len(t) == 9
        for k, c in cols.items():
            assert len(t[c].split(".")[1]) == PRINTED[k]
            if k in r["metrics"]:
                assert float(t[c]) == r["metrics"][k], (r["id"], k, t[c], r["metrics"][k])
                checked += 1
    assert checked >= 110


def test_arena_eval_qscale_recon_and_psnr(cc, clips, tmp_path):
    """--qscale 1 is the default codec; a coarser --qscale spends fewer bits;
    --dump-recon writes the decoded clip, whose `arena psnr` against the
    source is the eval's PSNR; --clips restricts and orders the clip set."""
    from csfunnel.arena import main
    kat = json.load(open(os.path.join(GOLDEN, "halloffame_kats.json")))["records"][0]
    corp = _corpus_subset(clips, str(tmp_path / "corpus"), LEGACY5)
    pop = tmp_path / "pop.txt"
    pop.write_text(f"{kat['id']} 0 0 0 0 " + " ".join("%.4f" % x for x in kat["g"]) + "\n")
    sub = tmp_path / "clips.txt"
    sub.write_text("# subset\ngrid\nnot-a-clip\n")

    def run(*extra):
        out = tmp_path / "r.csv"
        assert main(["eval", "--corpus", corp, "--pop", str(pop), "--out", str(out), "--clips", str(sub)]
                    + list(extra)) == 0
        t = out.read_text().split(",")
        return [float(x) for x in t[1:3] + t[4:]]

    base = run()
    assert run("--qscale", "1.0") == base
    coarse = run("--qscale", "2.0")
    assert coarse[1] < base[1] and coarse[0] < base[0]
    rd = tmp_path / "recon"
    assert run("--dump-recon", str(rd)) == base
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(os.path.dirname(GOLDEN), "..", "codec-arena_amd", "bin", "arena"),
                        "psnr", os.path.join(corp, "grid.yuv"), str(rd / "grid.recon.yuv"), "320", "192", "24"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    y, u, v = (float(x) for x in r.stdout.split())
    assert abs(y - base[0]) <= 0.005 and abs(u - base[3]) <= 0.005 and abs(v - base[4]) <= 0.005


def test_arena_eval_errors(cc, clips, tmp_path, capfd):
    from csfunnel.arena import main
    assert main(["eval", "--corpus", str(tmp_path / "none"), "--pop", "x", "--out", str(tmp_path / "o")]) == 1
    assert "error: FileNotFound" in capfd.readouterr().err
    corp = _corpus_subset(clips, str(tmp_path / "corpus"), ("drift",))
    sub = tmp_path / "clips.txt"
    sub.write_text("nothing\n")
    assert main(["eval", "--corpus", corp, "--pop", "x", "--clips", str(sub)]) == 1
    assert "error: NoClipsMatched" in capfd.readouterr().err
    with open(os.path.join(corp, "corpus.txt"), "a") as f:
        f.write("short 320 192\n")
    assert main(["eval", "--corpus", corp, "--pop", "x"]) == 1
    assert "error: BadSpec" in capfd.readouterr().err
    assert main(["frobnicate"]) == 1 and main([]) == 1
