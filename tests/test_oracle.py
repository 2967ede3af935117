"""Pins the CPU oracle of the screen against the reference's own known answers,
textbook vectors and properties that the mathematics fixes.

Reference tests and facts used (codec-arena snapshot):
  src/screen.zig:345-377   Model.fromCompact layout facts (v1/v2/v3)
  src/screen.zig:379-405   score properties (baseq monotone, v0.4 genes neutral)
  src/genome.zig:296-312   parse of v1 and v2 lines
  src/genome.zig:116-188   token-count table (incl. the v6 behaviour)
  splitmix64               Steele, Lea & Flood (OOPSLA 2014); Vigna splitmix64.c
"""
import os

import numpy as np
import pytest

GAMMA = 0x9E3779B97F4A7C15
M64 = (1 << 64) - 1


# ------------------------------------------------------------------ RNG

def test_splitmix64_known_answers(oracle):
    # first outputs of splitmix64 from state 0 (published test vector)
    out = oracle.splitmix(0, 3)
    assert [int(x) for x in out] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_splitmix64_counter_form(oracle):
    # out(n) = mix(seed + n*gamma): the n-th draw depends only on n
    seed = 0x1234_5678_9ABC_DEF0
    seq = oracle.splitmix(seed, 1000)
    for n in (1, 2, 17, 999, 1000):
        one = oracle.splitmix((seed + (n - 1) * GAMMA) & M64, 1)
        assert int(one[0]) == int(seq[n - 1])


def test_stream_seeds_seed42(oracle):
    # src/screen.zig:284: seed ^ (gamma * (t+1)) with wrapping multiply
    want = [0x9E3779B97F4A7C3F, 0x3C6EF372FE94F800, 0xDAA66D2C7DDF7415, 0x78DDE6E5FD29F07E]
    assert [oracle.stream_seed(42, t) for t in range(4)] == want
    for t in range(4):
        assert oracle.stream_seed(42, t) == 42 ^ ((GAMMA * (t + 1)) & M64)


# ---------------------------------------------------------------- model

def test_from_compact_reference_facts(oracle):
    m1 = oracle.model_from_compact(np.arange(153, dtype=np.float32))
    assert m1.w0 == 0 and m1.lin[0] == 1 and m1.lin[16] == 0
    assert m1.quad[0] == 17 and m1.quad[16] == 18 and m1.quad[1] == 0
    # src/screen.zig:357-358 asserts quad[33] == 20 ("17+T(2)+1 = 20"), but
    # T(2) = 3 and the fill loop of src/screen.zig:88-93 stores w[21] there;
    # that test is unreachable in the reference build.  The code is followed.
    assert m1.quad[33] == 21 and m1.quad[32] == 20
    m2 = oracle.model_from_compact(np.arange(159, dtype=np.float32))
    assert m2.lin[21] == 22 and m2.quad[0] == 23
    m3 = oracle.model_from_compact(np.arange(161, dtype=np.float32))
    assert m3.lin[16] == 17 and m3.lin[17] == 18 and m3.lin[18] == 0 and m3.lin[23] == 19
    assert m3.quad[0] == 25


@pytest.mark.parametrize("n,lin_extra,regs_at,quad_at", [
    (153, 0, None, 17), (159, 6, None, 23), (161, 2, 19, 25), (164, 5, 22, 28), (165, 6, 23, 29),
    (166, 7, 24, 30)])
def test_from_compact_all_widths(oracle, n, lin_extra, regs_at, quad_at):
    w = np.arange(n, dtype=np.float32) + 1000
    m = oracle.model_from_compact(w)
    lin = np.array(m.lin[:])
    assert np.array_equal(lin[:16], w[1:17])
    assert np.array_equal(lin[16:16 + lin_extra], w[17:17 + lin_extra])
    if regs_at is not None:
        assert np.array_equal(lin[23:29], w[regs_at:regs_at + 6])
    quad = np.array(m.quad[:]).reshape(16, 16)
    tri = [quad[i, j] for i in range(16) for j in range(i + 1)]
    assert np.array_equal(np.array(tri), w[quad_at:quad_at + 136])
    assert np.all(np.triu(quad, 1) == 0)
    assert quad_at + 136 == n


def test_from_compact_rejects_bad_length(oracle):
    with pytest.raises(ValueError):
        oracle.model_from_compact(np.zeros(160, dtype=np.float32))


# ---------------------------------------------------------------- parse

V1 = "1 0 0 0 7 0.1 0.9 0.2 0.3 0.4 0.5 0.6 0.7 0.8 0.9 0.05 0.5 0.5 0.5 0.11 0.22"
V2 = ("2 0 0 0 7 0.1 0.9 0.2 0.3 0.4 0.5 0.6 0.7 0.8 0.9 0.05 0.5 0.5 0.5 0.11 0.22"
      " 0.9 0.1 0.5 0.5 0.5 0.7")


def f32(x):
    return float(np.float32(x))


def test_parse_v1_v2_reference_kat(oracle):
    rc, g1 = oracle.parse_line(V1)
    assert rc == 1 and g1.id == 1 and g1.tribe == 7
    assert g1.reg[0] == 0.5 and g1.reg[5] == 0.5 and g1.g[16] == 0 and g1.g[17] == 0
    assert g1.g[0] == f32(0.1) and g1.g[15] == f32(0.22)
    rc, g2 = oracle.parse_line(V2)
    assert rc == 1
    assert g2.reg[0] == f32(0.9) and g2.reg[1] == f32(0.1) and g2.reg[5] == f32(0.7) and g2.g[16] == 0


def _line(n):
    head = "5 3 1 2 9"
    vals = [f"{(k + 1) / 100:.2f}" for k in range(n - 5)]
    return head + " " + " ".join(vals)


@pytest.mark.parametrize("n", [21, 27, 29, 32, 33, 34])
def test_parse_token_table(oracle, n):
    rc, g = oracle.parse_line(_line(n))
    assert rc == 1
    t = [f32((k + 1) / 100) for k in range(n - 5)]  # token 5+k -> t[k]
    G = np.array(g.g[:])
    R = np.array(g.reg[:])
    assert np.array_equal(G[:16], np.array(t[:16], dtype=np.float32))
    want16 = {29: t[16:18], 32: t[16:18], 33: t[16:18]}.get(n, [0, 0])
    assert list(G[16:18]) == list(np.float32(want16))
    want18 = {32: t[18:21], 33: t[18:21]}.get(n, [0, 0, 0])
    assert list(G[18:21]) == list(np.float32(want18))
    assert G[21] == (t[21] if n == 33 else 0)
    assert G[22] == (t[22] if n == 34 else 0)  # v6: token 27
    rstart = {27: 16, 29: 18, 32: 21, 33: 22, 34: 23}.get(n)
    if rstart is None:
        assert np.all(R == 0.5)
    else:
        assert np.array_equal(R, np.array(t[rstart:rstart + 6], dtype=np.float32))


def test_parse_v6_line_drops_genes_16_to_21(oracle, golden):
    # every line of the reference population is v6 (34 tokens)
    with open(os.path.join(golden, "seed_pop.txt")) as f:
        line = f.readline().strip()
    toks = line.split()
    assert len(toks) == 34
    rc, g = oracle.parse_line(line)
    assert rc == 1
    assert list(g.g[16:22]) == [0.0] * 6
    assert g.g[22] == f32(float(toks[27]))
    assert list(g.reg[:]) == [f32(float(x)) for x in toks[28:34]]


def test_parse_skips_and_errors(oracle):
    assert oracle.parse_line("1 2 3")[0] == 0
    assert oracle.parse_line("# " + V1[2:])[0] == 0
    assert oracle.parse_line(V1 + " 0.1" * 14)[0] == -5  # 35 tokens -> BadLine
    assert oracle.parse_line(V1.replace("0.9", "x9", 1))[0] < 0
    assert oracle.parse_line("1_0" + V1[1:])[1].id == 10  # Zig digit separators
    assert oracle.parse_line("-1" + V1[1:])[0] < 0


# ---------------------------------------------------------------- score

def features153(g):
    """brain/evolve.py:211-220 on the 16 legacy genes, FP64."""
    v = [1.0] + list(g[:16])
    for i in range(16):
        for j in range(i + 1):
            v.append(g[i] * g[j])
    return np.array(v)


def test_score_matches_fp64_polynomial(oracle, golden):
    with open(os.path.join(golden, "c0_model.txt")) as f:
        toks = f.read().split()
    w = np.array([np.float32(x) for x in toks[1:]], dtype=np.float32)
    m = oracle.model_from_compact(w)
    rng = np.random.default_rng(5)
    for _ in range(200):
        g = rng.random(29).astype(np.float32)
        s32 = oracle.score(m, g)
        phi = features153(g.astype(np.float64))
        # src/screen.zig:100 splats w0 into all 16 lanes before the lane sum,
        # so the intercept enters the score 16 times (a constant offset).
        s64 = float(phi @ w.astype(np.float64)) + 15.0 * float(w[0])
        bound = 40 * 2.0 ** -24 * (float(np.abs(phi * w.astype(np.float64)).sum()) + 15 * abs(float(w[0])))
        assert abs(s32 - s64) <= bound, (s32, s64, bound)


def test_score_reference_properties(oracle):
    w = np.zeros(153, dtype=np.float32)
    w[0] = 1.0
    w[1 + 3] = 10.0
    m = oracle.model_from_compact(w)
    ga = np.full(29, 0.5, np.float32)
    gb = ga.copy()
    gb[3] = 0.9
    assert oracle.score(m, gb) > oracle.score(m, ga)
    gc = ga.copy()
    gc[16] = 0.95
    gc[17] = 0.9
    assert oracle.score(m, ga) == oracle.score(m, gc)


def test_score_linear_tail_166(oracle):
    # the 166 layout scores genes 16..28 linearly (screen.zig:108-110)
    w = np.zeros(166, dtype=np.float32)
    w[17 + 6] = 2.0  # sb16 -> flat 22
    w[24 + 5] = 4.0  # last regulator -> flat 28
    m = oracle.model_from_compact(w)
    g = np.zeros(29, np.float32)
    g[22] = 0.25
    g[28] = 0.5
    assert oracle.score(m, g) == 2.5


# ------------------------------------------------------------ synthesis

def _elites(oracle, golden, name="c0_elites.txt"):
    return oracle.load_elites(os.path.join(golden, name))


def test_crossover_without_mutation_picks_parent_genes(oracle, golden):
    el = _elites(oracle, golden)
    E = np.array([list(e.g[:]) + list(e.reg[:]) for e in el], dtype=np.float32)
    start = 0
    for _ in range(300):
        g, L, ia, ib = oracle.synth(el, 77, start, 0.0, 0.0, 0.0)
        assert L == 62
        assert np.all((g == E[ia]) | (g == E[ib]))
        start += L


def test_immigrants_and_jumps_draw_counts(oracle, golden):
    el = _elites(oracle, golden)
    start = 0
    for _ in range(200):
        g, L, _, _ = oracle.synth(el, 99, start, 0.3, 0.0, 1.0)  # always immigrant
        assert L == 33
        m = g * 16777216.0
        assert np.all(m == np.floor(m)) and np.all(g < 1) and np.all(g >= 0)
        start += L
    seen = set()
    for _ in range(200):
        g, L, _, _ = oracle.synth(el, 99, start, 0.3, 1.0, 0.0)  # always jump
        seen.add(L)
        start += L
    assert seen == {64}


def test_mutation_bounded_by_sigma(oracle, golden):
    el = _elites(oracle, golden)
    E = np.array([list(e.g[:]) + list(e.reg[:]) for e in el], dtype=np.float32)
    sigma = 0.05
    start = 0
    for _ in range(300):
        g, L, ia, ib = oracle.synth(el, 3, start, sigma, 0.0, 0.0)
        d = np.minimum(np.abs(g - E[ia]), np.abs(g - E[ib]))
        assert np.all(d <= sigma * (1 + 1e-6))
        assert np.all((g >= 0) & (g <= 1))
        start += L


def _enumerate(oracle, el, seed, lam, sigma, jump_p, imm_p):
    """Every candidate of one stream by chaining single-candidate synthesis."""
    out = []
    start = 0
    for i in range(lam):
        g, L, ia, ib = oracle.synth(el, seed, start, sigma, jump_p, imm_p)
        out.append((g, start, ia, ib))
        start += L
    return out


def test_screen_topk_equals_brute_force_sort(oracle, golden):
    el = _elites(oracle, golden)
    m = oracle.load_model(os.path.join(golden, "c0_model.txt"))
    lam, keep, T = 4000, 37, 4
    surv, screens = oracle.screen(m, el, lam, keep, seed=11, threads=T, sigma=0.12, jump_p=0.05, imm_p=0.05)
    assert screens == lam
    allc = []
    per = lam // T
    for t in range(T):
        my = per if t < T - 1 else lam - per * (T - 1)
        for k, (g, start, ia, ib) in enumerate(_enumerate(oracle, el, oracle.stream_seed(11, t), my, 0.12, 0.05,
                                                          0.05)):
            allc.append((oracle.score(m, g), t, k, start, g))
    allc.sort(key=lambda r: -r[0])
    assert len(set(r[0] for r in allc[: keep + 1])) == keep + 1  # no ties at the cut
    for s, want in zip(surv, allc[:keep]):
        assert s["score"] == np.float32(want[0])
        assert (s["stream"], s["index"], s["start"]) == (want[1], want[2], want[3])
        assert np.array_equal(s["g"], want[4])


def test_screen_ties_keep_first_k_in_stream_order(oracle, golden):
    # all-zero model: every score is 0; the heaps keep the first `keep`
    # candidates of each stream in insertion order, the stable merge is t-major
    el = _elites(oracle, golden)
    m = oracle.model_from_compact(np.zeros(153, np.float32))
    surv, _ = oracle.screen(m, el, 1000, 300, seed=5, threads=4)
    assert list(surv["stream"]) == [0] * 250 + [1] * 50
    assert list(surv["index"]) == list(range(250)) + list(range(50))


def test_partition_law_T4_equals_four_T1_runs(oracle, golden):
    el = _elites(oracle, golden)
    m = oracle.load_model(os.path.join(golden, "c0_model.txt"))
    lam, keep, seed = 20000, 50, 123
    s4, _ = oracle.screen(m, el, lam, keep, seed, threads=4)
    parts = []
    for t in range(4):
        st = oracle.stream_seed(seed, t)
        s1, _ = oracle.screen(m, el, lam // 4, keep, st ^ GAMMA, threads=1)
        assert oracle.stream_seed(st ^ GAMMA, 0) == st
        parts.extend((float(r["score"]), t, tuple(r["g"])) for r in s1)
    parts.sort(key=lambda r: -r[0])
    assert [float(x) for x in s4["score"]] == [p[0] for p in parts[:keep]]


# ------------------------------------------------------------- printing

@pytest.mark.parametrize("v,s", [(0.5, "0.5000"), (1.0, "1.0000"), (0.0, "0.0000"), (0.25, "0.2500"),
                                 (0.03125, "0.0313"), (0.12345, "0.1235"), (5e-05, "0.0001"),
                                 (4e-05, "0.0000"), (0.99996, "1.0000"), (0.1, "0.1000")])
def test_fmt_d4_known(oracle, v, s):
    assert oracle.fmt(v) == s


def test_fmt_d4_matches_decimal_rounding_of_shortest_repr(oracle):
    # Zig rounds the shortest f32 digits half-up; numpy gives the same digits
    rng = np.random.default_rng(1)
    for v in rng.random(3000).astype(np.float32):
        short = np.format_float_positional(v, unique=True, trim="-")
        from decimal import Decimal, ROUND_HALF_UP
        want = str(Decimal(short).quantize(Decimal("0.0001"), rounding=ROUND_HALF_UP))
        if Decimal(short) < Decimal("0.0001"):
            continue  # below 10^-4 Zig rounds on the leading digit, not at the 4th decimal
        assert oracle.fmt(float(v)) == want, (repr(v), short)
