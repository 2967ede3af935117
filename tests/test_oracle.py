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


def _flat_pop(tmp_path, ne=8, value="0.5000"):
    """ne 33-token (v5) lines with every gene and regulator = value: genes
    0..21 and the regulators parse to it, gene 22 to 0 (src/genome.zig:116-188)."""
    p = tmp_path / f"flat_{ne}.txt"
    p.write_text("".join(f"{i} 0 0 0 {i % 12} " + " ".join([value] * 28) + "\n" for i in range(ne)))
    return str(p)


def _draws(oracle, seed, c, n):
    """out(c + 1) .. out(c + n) of the stream with state seed (splitmix64)."""
    return oracle.splitmix((seed + c * GAMMA) % 2**64, n)


PARENT_HALF = [k for k in range(29) if k != 22]  # genes whose parents are 0.5 in _flat_pop


def test_mutation_closed_form_at_sigma_half(oracle, tmp_path):
    # parent 0.5, sigma 0.5: 0.5 + ((u - 0.5) * 2) * 0.5 = u exactly, so every
    # such gene is the mutation unit itself, on the 2^-24 lattice, and the
    # offsets g - 0.5 fill [-sigma, sigma).  A dropped "* 2" halves the range,
    # a dropped "- 0.5" pushes every gene to >= 0.5 (and clamps half of them).
    el = oracle.load_elites(_flat_pop(tmp_path))
    seed, start, offs = 12345, 0, []
    for _ in range(3000):
        g, L, ia, ib = oracle.synth(el, seed, start, 0.5, 0.0, 0.0)
        assert L == 62
        d = _draws(oracle, seed, start, L)
        u = (d[4:62:2] >> np.uint64(40)).astype(np.float64) * 2.0**-24  # draws c+5, c+7, ..., c+61
        assert np.array_equal(g[PARENT_HALF].astype(np.float64), u[PARENT_HALF])
        offs.append(g[PARENT_HALF].astype(np.float64) - 0.5)
        start += L
    offs = np.concatenate(offs)
    assert offs.min() >= -0.5 and offs.max() < 0.5
    assert offs.min() <= -0.5 + 1e-3 * 0.5 and offs.max() >= 0.5 - 1e-3 * 0.5
    assert abs(offs.mean()) < 0.01


@pytest.mark.parametrize("k", [3, 7, 12])
def test_mutation_range_at_power_of_two_sigma(oracle, tmp_path, k):
    # sigma = 2^-k: offsets stay inside [-sigma, sigma] (one rounding of the
    # sum at most) and reach within 1e-3 sigma of both ends
    el = oracle.load_elites(_flat_pop(tmp_path))
    sigma, start, offs = 2.0**-k, 0, []
    for _ in range(4000):
        g, L, _, _ = oracle.synth(el, 777 + k, start, sigma, 0.0, 0.0)
        offs.append(g[PARENT_HALF].astype(np.float64) - 0.5)
        start += L
    offs = np.concatenate(offs)
    ulp = 2.0**-25
    assert offs.min() >= -sigma - ulp and offs.max() <= sigma + ulp
    assert offs.min() <= -sigma + 1e-3 * sigma and offs.max() >= sigma - 1e-3 * sigma


@pytest.mark.parametrize("ne", [128, 96, 7])
def test_elite_indices_are_high_words_mod_ne(oracle, tmp_path, ne):
    # ea = elites[(out(c+1) >> 32) % ne], eb = elites[(out(c+2) >> 32) % ne]
    el = oracle.load_elites(_flat_pop(tmp_path, ne))
    seed, start = 0xDEADBEEF, 0
    for _ in range(2000):
        _, L, ia, ib = oracle.synth(el, seed, start, 0.1, 0.02, 0.01)
        d = _draws(oracle, seed, start, 2)
        assert (ia, ib) == (int(d[0] >> np.uint64(32)) % ne, int(d[1] >> np.uint64(32)) % ne)
        start += L


def test_immigrant_and_jump_tests_read_their_draws(oracle, tmp_path):
    # immigrant iff unit(out(c+3)) < imm_p; the jump test is out(c+33) for an
    # immigrant and out(c+62) otherwise, with strict "<"
    el = oracle.load_elites(_flat_pop(tmp_path))
    seed, start = 4242, 0
    kinds = set()
    for _ in range(3000):
        _, L, _, _ = oracle.synth(el, seed, start, 0.1, 0.5, 0.5)
        d = _draws(oracle, seed, start, 64)
        unit = lambda x: float(int(x) >> 40) * 2.0**-24  # noqa: E731
        imm = unit(d[2]) < 0.5
        jump = unit(d[32 if imm else 61]) < 0.5
        assert L == (33 if imm else 62) + (2 if jump else 0)
        kinds.add(L)
        start += L
    assert kinds == {33, 35, 62, 64}


def test_event_tests_are_strict_at_the_rate(oracle, tmp_path):
    # a rate equal to the drawn unit is not an event (unit < p), the next f32
    # above it is: immigrant test at out(c+3), jump test at out(c+62)
    el = oracle.load_elites(_flat_pop(tmp_path))
    unit = lambda x: np.float32(float(int(x) >> 40) * 2.0**-24)  # noqa: E731
    for seed, c in ((1, 0), (2, 1000), (3, 77)):
        d = _draws(oracle, seed, c, 64)
        ui, uj = unit(d[2]), unit(d[61])
        above = lambda u: float(np.nextafter(u, np.float32(2)))  # noqa: E731
        assert oracle.synth(el, seed, c, 0.1, 0.0, float(ui))[1] == 62
        assert oracle.synth(el, seed, c, 0.1, 0.0, above(ui))[1] == 33
        assert oracle.synth(el, seed, c, 0.1, float(uj), 0.0)[1] == 62
        assert oracle.synth(el, seed, c, 0.1, above(uj), 0.0)[1] == 64


def test_jump_index_uniform_over_29_and_value_on_lattice(oracle, tmp_path):
    # sigma 0 and every parent gene 0.5 except gene 22 (0): the jumped gene is
    # the one that moved; its index is (out(c+63) >> 32) % 29, uniform over
    # the 29 genes (chi-square, 28 degrees of freedom), and its value is
    # unit(out(c+64)) on the 2^-24 lattice
    el = oracle.load_elites(_flat_pop(tmp_path))
    base = np.full(29, 0.5, np.float32)
    base[22] = 0.0
    seed, start = 99, 0
    counts = np.zeros(29, int)
    for _ in range(29 * 400):
        g, L, _, _ = oracle.synth(el, seed, start, 0.0, 1.0, 0.0)
        assert L == 64
        d = _draws(oracle, seed, start, 64)
        idx = int(d[62] >> np.uint64(32)) % 29
        val = float(int(d[63]) >> 40) * 2.0**-24
        expect = base.copy()
        expect[idx] = val
        assert np.array_equal(g, expect)
        assert float(g[idx]) * 2**24 == int(float(g[idx]) * 2**24)
        counts[idx] += 1
        start += L
    chi2 = float(((counts - 400.0) ** 2 / 400.0).sum())
    assert chi2 < 80.0, (chi2, counts)  # p < 1e-6 for 28 degrees of freedom


def test_clamp01_special_values(oracle):
    # NaN -> 1 and -0 -> +0 (the constants win ties and unordered compares)
    cases = [(float("nan"), 1.0), (-0.0, 0.0), (2.0, 1.0), (-1.0, 0.0), (float("inf"), 1.0),
             (float("-inf"), 0.0), (0.5, 0.5), (1e-45, 1e-45), (1.0, 1.0), (0.0, 0.0)]
    for x, want in cases:
        got = np.float32(oracle.L.or_clamp01(np.float32(x)))
        assert got.view(np.uint32) == np.float32(want).view(np.uint32), (x, got)


def test_simd_score_form_equals_scalar(oracle, golden):
    rng = np.random.default_rng(17)
    for n in (153, 159, 161, 164, 165, 166):
        m = oracle.model_from_compact((rng.standard_normal(n) / np.sqrt(n)).astype(np.float32))
        for _ in range(500):
            g = rng.random(29).astype(np.float32)
            g[rng.random(29) < 0.1] = 0.0
            a = np.float32(oracle.score(m, g))
            b = np.float32(oracle.score_simd(m, g))
            assert a.view(np.uint32) == b.view(np.uint32)


def _score_first_product_order(m, g):
    """The device scorer's float32 operation order: each lane's quadratic sum
    and the linear tail start at their first product instead of at +0, and
    the score gets a final + 0.0 (numpy float32 arithmetic: round to nearest,
    no fused multiply-add)."""
    f = np.float32
    w0, lin = f(m.w0), np.array(m.lin[:], np.float32)
    quad = np.array(m.quad[:], np.float32).reshape(16, 16)
    S = None
    for j in range(16):
        acc = f(w0 + f(lin[j] * g[j]))
        s = f(quad[j, j] * f(g[j] * g[j]))
        for i in range(j + 1, 16):
            s = f(s + f(quad[i, j] * f(g[j] * g[i])))
        ln = f(acc + s)
        S = ln if j == 0 else f(S + ln)
    r = f(lin[16] * g[16])
    for i in range(17, 29):
        r = f(r + f(lin[i] * g[i]))
    return f(f(S + r) + f(0.0))


def test_score_from_first_product_equals_reference_order(oracle):
    # The scorer drops the reference's +0 starts of its sums (17 adds per
    # candidate) and adds +0 once at the end.  The two orders may differ only
    # in the sign of a zero partial result, and the reference's score is
    # never -0, so the scores must agree bit for bit.  Inputs full of signed
    # zeros (weights, genes, the intercept) are where a difference would show.
    rng = np.random.default_rng(29)
    palette = np.array([0.0, -0.0, 1.0, -1.0, 0.5, -2.0, 2.0 ** -149, -(2.0 ** -149)], np.float32)
    zero_scores = 0
    for case in range(400):
        n = (153, 166)[(case // 8) % 2]
        if case % 4 == 3:  # signed zeros only: every score is a zero
            w = palette[rng.integers(0, 2, n)]
        elif case % 2:
            w = palette[rng.integers(0, palette.size, n)]
        else:
            w = (rng.standard_normal(n) * (rng.random(n) < 0.3)).astype(np.float32)
            w[rng.random(n) < 0.3] = -0.0
        w[0] = [0.0, -0.0, 1.0, [0.0, -0.0][(case // 4) % 2]][case % 4]
        m = oracle.model_from_compact(w.astype(np.float32))
        for _ in range(8):
            g = rng.random(29).astype(np.float32)
            g[rng.random(29) < 0.5] = 0.0
            g[rng.random(29) < 0.3] = -0.0
            want = np.float32(oracle.score(m, g))
            got = _score_first_product_order(m, g)
            assert got.view(np.uint32) == want.view(np.uint32), (case, want, got)
            if want == 0:
                zero_scores += 1
                assert not np.signbit(want)
    assert zero_scores >= 800  # the signed-zero cases were reached
    # every weight -0.0 and genes >= +0: every product is -0.0, so sums that
    # start at their first product are -0.0 and only the final + 0.0 gives
    # the reference's +0
    # (width 166: the linear tail over genes 16..28 comes from the weights
    # too; at width 153 it is +0 and would restore the sign by itself)
    m = oracle.model_from_compact(np.full(166, -0.0, np.float32))
    for _ in range(20):
        g = rng.random(29).astype(np.float32)
        g[rng.random(29) < 0.3] = 0.0
        want = np.float32(oracle.score(m, g))
        assert want == 0 and not np.signbit(want)
        assert _score_first_product_order(m, g).view(np.uint32) == want.view(np.uint32)


@pytest.mark.parametrize("lam,T,sigma,jump,imm", [(1 << 20, 4, 0.12, 0.01, 0.01), (1 << 18, 1, 0.12, 0.01, 0.01),
                                                 (200000, 3, 0.5, 0.5, 0.5), (100000, 4, 0.0, 0.0, 0.0)])
def test_simd_screen_equals_scalar_screen(oracle, golden, lam, T, sigma, jump, imm):
    m = oracle.load_model(os.path.join(golden, "c0_model.txt"))
    el = _elites(oracle, golden)
    args = (m, el, lam, 256, 5, T, float(np.float32(sigma)), float(np.float32(jump)), float(np.float32(imm)))
    a, sa = oracle.screen(*args)
    b, sb = oracle.screen(*args, simd=True)
    assert sa == sb == lam
    assert a.tobytes() == b.tobytes()


@pytest.mark.parametrize("case", ["c0", "c0_T1", "rates", "zero_model", "coarse", "keep_big"])
def test_chunked_screen_equals_sequential(oracle, golden, tmp_path, case):
    # the chunked form (per-chunk heaps, replay of what they took) against the
    # plain per-stream worker loop, including exact ties straddling the cut
    m = oracle.load_model(os.path.join(golden, "c0_model.txt"))
    el = _elites(oracle, golden)
    f = lambda x: float(np.float32(x))  # noqa: E731
    args = dict(lam=1 << 20, keep=256, seed=9, threads=4, sigma=f(0.12), jump_p=f(0.01), imm_p=f(0.01))
    if case == "c0_T1":
        args.update(threads=1, lam=1 << 19)
    elif case == "rates":
        args.update(lam=300000, threads=3, sigma=f(0.5), jump_p=f(0.5), imm_p=f(0.5))
    elif case == "zero_model":
        m = oracle.model_from_compact(np.zeros(153, np.float32))
    elif case == "coarse":
        rng = np.random.default_rng(8)
        m = oracle.model_from_compact(rng.integers(-2, 3, 153).astype(np.float32))
        pop = tmp_path / "coarse.txt"
        pop.write_text("".join(f"{i} 0 0 0 {i} " + " ".join(str(x) for x in rng.integers(0, 2, 16)) + "\n"
                               for i in range(8)))
        el = oracle.load_elites(str(pop))
        args.update(lam=400000, keep=500, sigma=0.0, jump_p=0.0, imm_p=0.0)
    elif case == "keep_big":
        args.update(lam=50000, keep=20000)
    ref, rs = oracle.screen(m, el, **args)
    for chunks in (1, 7, 64):
        got, gs = oracle.screen_chunked(m, el, chunks=chunks, **args)
        assert gs == rs == args["lam"]
        assert got.tobytes() == ref.tobytes(), chunks


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
