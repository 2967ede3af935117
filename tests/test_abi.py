"""CPU checks of the C-ABI library: it loads, exports every entry point that
include/csfunnel.h declares, and its host-side parsing/format code agrees with
the oracle.  No device calls are made here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from csfunnel import _abi, screen as scr

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(REPO, "include", "csfunnel.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cs_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = _abi.lib()
    names = header_functions()
    assert len(names) >= 14
    for n in names:
        assert hasattr(L, n), n
        assert n in _abi.SIGNATURES, n


def test_struct_sizes_match_header():
    def pad8(n):
        return (n + 7) // 8 * 8
    assert C.sizeof(_abi.Genome) == pad8(8 + 4 + 4 + 8 + 8 + 23 * 4 + 6 * 4)
    assert C.sizeof(_abi.Survivor) == pad8(4 + 29 * 4 + 8 + 8 + 4 + 4 + 8 + 8)
    assert C.sizeof(_abi.ScreenCfg) == 48


def _genome_tuple(g):
    return (g.id, g.gen, g.tribe, g.pa, g.pb, tuple(np.float32(g.g[:])), tuple(np.float32(g.reg[:])))


def test_parse_lines_match_oracle(oracle, golden):
    lines = []
    for name in ("seed_pop.txt", "c0_elites.txt"):
        lines += open(os.path.join(golden, name)).read().splitlines()
    base = lines[0].split()
    for n in (21, 27, 29, 32, 33, 34, 35, 20):
        toks = (base * 2)[:n]
        lines.append(" ".join(toks))
    lines += ["# comment", "", "1 0 0 0 1" + " 0.5" * 16 + " nan", "1 0 0 0 1" + " 0.5" * 15 + " 1e-3"]
    for line in lines:
        rc_o, go = oracle.parse_line(line)
        g = _abi.Genome()
        b = line.encode()
        rc = _abi.lib().cs_parse_pop_line(b, len(b), C.byref(g))
        assert rc == rc_o or (rc < 0 and rc_o < 0), line
        if rc == 1:
            a, o = _genome_tuple(g), _genome_tuple(go)
            assert a[:5] == o[:5]
            assert np.array_equal(np.array(a[5]), np.array(o[5]), equal_nan=True)
            assert np.array_equal(np.array(a[6]), np.array(o[6]), equal_nan=True)


def test_load_elites_and_models_match_oracle(oracle, golden):
    for name in ("seed_pop.txt", "c0_elites.txt", "c1_elites.txt"):
        p = os.path.join(golden, name)
        mine = scr.load_elites(p)
        ref = oracle.load_elites(p)
        assert len(mine) == len(ref)
        for a, b in zip(mine, ref):
            assert _genome_tuple(a) == _genome_tuple(b)
    for name in ("c0_model.txt", "surrogate_161.txt"):
        p = os.path.join(golden, name)
        m = scr.load_model(p).raw
        o = oracle.load_model(p)
        assert bytes(m) == bytes(o)
    rng = np.random.default_rng(3)
    for n in (153, 159, 161, 164, 165, 166):
        w = rng.standard_normal(n).astype(np.float32)
        assert bytes(scr.Model.from_compact(w).raw) == bytes(oracle.model_from_compact(w))
    with pytest.raises(_abi.CsError):
        scr.Model.from_compact(np.zeros(152, np.float32))


def test_format_gene_two_implementations_agree(oracle):
    rng = np.random.default_rng(9)
    vals = list(rng.random(20000).astype(np.float32))
    vals += list((rng.random(2000) * 1e-3).astype(np.float32))
    vals += list(np.float32(np.arange(0, 2 ** 16)) * np.float32(2 ** -16))[::7]
    vals += [np.float32(x) for x in (0.0, -0.0, 1.0, 0.5, 5e-5, 4.99e-5, 5e-6, 9.7e-6, 1e-7, 0.99995, 0.10005,
                                     0.00015, 2.5, 12.34565, 1e-30)]
    for v in vals:
        assert scr.format_gene(float(v)) == oracle.fmt(float(v)), repr(v)


def test_writers_match_oracle(oracle, tmp_path):
    rng = np.random.default_rng(4)
    n = 50
    s = np.zeros(n, dtype=scr.SURVIVOR_DTYPE)
    s["score"] = rng.standard_normal(n)
    s["g"] = rng.random((n, 29)).astype(np.float32)
    s["g"][:, 0] = 0.0
    s["g"][:, 1] = 1.0
    s["pa"] = rng.integers(0, 2 ** 40, n)
    s["pb"] = rng.integers(0, 2 ** 40, n)
    s["tribe"] = rng.integers(0, 12, n)
    a, b = tmp_path / "a.txt", tmp_path / "b.txt"
    scr.write_promoted(str(a), s)
    oracle.write_promoted(str(b), s)
    assert a.read_bytes() == b.read_bytes()
    assert len(a.read_text().splitlines()[0].split()) == 29
    scr.write_count(str(a), 2650000000)
    oracle.write_count(str(b), 2650000000)
    assert a.read_bytes() == b.read_bytes() == b"2650000000\n"


def test_parse_scalars_zig_rules():
    assert scr.parse_u64("1_000") == 1000
    assert scr.parse_u64("+7") == 7
    for bad in ("", "-1", "_1", "1_", "0x10", "18446744073709551616"):
        with pytest.raises(_abi.CsError):
            scr.parse_u64(bad)
    assert scr.parse_f32("0.1200") == float(np.float32(0.12))
    with pytest.raises(_abi.CsError):
        scr.parse_f32("0.1x")


def _write_pop(path, lines):
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def test_screen_cmd_host_errors(tmp_path, golden, capsys):
    m = os.path.join(golden, "c0_model.txt")
    e = os.path.join(golden, "c0_elites.txt")
    out, cnt = str(tmp_path / "p.txt"), str(tmp_path / "c.txt")
    # trailing flag without a value
    assert scr.screen_cmd(["--model", m, "--keep"]) == 1
    assert "BadArgs" in capsys.readouterr().err
    # bad model length
    bad = tmp_path / "bad.txt"
    bad.write_text("160\n" + " 0" * 160)
    assert scr.screen_cmd(["--model", str(bad), "--elites", e, "--out", out, "--count-out", cnt]) == 1
    assert "BadModel" in capsys.readouterr().err
    # fewer than two elites
    one = tmp_path / "one.txt"
    _write_pop(one, open(e).read().splitlines()[:1] + ["# c", "1 2 3"])
    assert scr.screen_cmd(["--model", m, "--elites", str(one), "--out", out, "--count-out", cnt]) == 1
    assert "NeedTwoElites" in capsys.readouterr().err
    # bad integer
    assert scr.screen_cmd(["--lambda", "12x", "--model", m]) == 1
    assert not os.path.exists(out) and not os.path.exists(cnt)


def test_caller_cfg_mirrors_evolve_call():
    """brain/evolve.py:368-371, 431-438 for pop 256, seed 42, shard 7."""
    cfg = scr.caller_cfg(2650000000, 256, 42, 7, 3)
    assert (cfg.lambda_, cfg.keep, cfg.seed, cfg.threads) == (2650000000, 256 - 32 - 32, 42 * 1000003 + 7 * 131 + 3, 4)
    assert cfg.sigma == float(np.float32(0.0965)) and "%.4f" % (0.12 * 0.93 ** 3) == "0.0965"
    assert cfg.jump_p == float(np.float32(0.02)) and cfg.imm_p == float(np.float32(0.01))
    assert scr.caller_cfg(10, 12, 1, 0, 0).keep == 12 - 2 - 2
    assert scr.caller_cfg(10, 12, 1, 0, 0, keep=5).keep == 5


def _evolve_promoted(path, ng=29):
    """brain/evolve.py:445-454 restated: lines of 5 + NG tokens, pa/pb/tribe
    as ints and the genes as floats; every other line is dropped."""
    out = []
    for line in open(path):
        t = line.split()
        if len(t) != 5 + ng:
            continue
        out.append({"pa": int(t[2]), "pb": int(t[3]), "tribe": int(t[4]), "g": [float(x) for x in t[5:]]})
    return out


def test_v6_writer_lines_are_kept_by_evolve(tmp_path):
    rng = np.random.default_rng(5)
    s = np.zeros(7, dtype=scr.SURVIVOR_DTYPE)
    s["score"] = rng.standard_normal(7)
    s["g"] = rng.random((7, 29)).astype(np.float32)
    s["g"][0, :3] = [0.0, 1.0, 5e-5]
    s["pa"], s["pb"], s["tribe"] = rng.integers(0, 10**12, 7), rng.integers(0, 10**12, 7), rng.integers(0, 12, 7)
    zig, v6 = str(tmp_path / "zig.txt"), str(tmp_path / "v6.txt")
    scr.write_promoted(zig, s)
    scr.write_promoted_v6(v6, s)
    assert _evolve_promoted(zig) == []  # R2: the reference's 29-token lines are all dropped
    kept = _evolve_promoted(v6)
    assert len(kept) == 7
    for r, row in zip(s, kept):
        assert (row["pa"], row["pb"], row["tribe"]) == (r["pa"], r["pb"], r["tribe"])
        assert row["g"] == [float(scr.format_gene(float(x))) for x in r["g"]]
    # the v6 line's first 29 tokens are the reference line, token for token
    for a, b in zip(open(zig), open(v6)):
        assert b.split()[:29] == a.split()
    # and the next generation reads them back through the v6 parse (R1: genes 16..21 read as 0)
    for line, r in zip(open(v6), s):
        g = scr.parse_line(line.rstrip("\n"))
        assert g is not None and g.pa == r["pa"] and list(g.g[16:22]) == [0.0] * 6
        assert np.float32(g.g[22]) == np.float32(float(scr.format_gene(float(r["g"][22]))))


def test_v6_format_environment_switch_is_checked(tmp_path, golden, monkeypatch, capsys):
    monkeypatch.setenv("CSFUNNEL_PROMOTED_FORMAT", "bogus")
    rc = scr.screen_cmd(["--model", os.path.join(golden, "c0_model.txt"), "--elites",
                         os.path.join(golden, "c0_elites.txt"), "--out", str(tmp_path / "p.txt")])
    assert rc == 1 and "CSFUNNEL_PROMOTED_FORMAT" in capsys.readouterr().err
    assert not (tmp_path / "p.txt").exists()
