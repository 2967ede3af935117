"""GPU parity of the quadratic screen: the HIP funnel (through the C ABI) must
return exactly the oracle's survivors -- scores, genes, lineage and stream
positions bit for bit, in the same order -- and screens_count == lambda."""
import os

import numpy as np
import pytest

from csfunnel import screen as scr
from csfunnel._abi import CsError

pytestmark = pytest.mark.gpu

GAMMA = 0x9E3779B97F4A7C15


@pytest.fixture(scope="module", params=["auto", "maps"])
def dev(request):
    s = scr.Screen(0)
    s.set_offset_mode(0 if request.param == "auto" else 1)
    s.mode = request.param
    yield s
    s.close()


def _c0(golden, name="c0_elites.txt", model="c0_model.txt"):
    return os.path.join(golden, model), os.path.join(golden, name)


def _same(gpu, ref):
    assert len(gpu) == len(ref)
    for f in ("score", "g", "pa", "pb", "tribe", "stream", "index", "start"):
        a, b = gpu[f], ref[f]
        if a.dtype.kind == "f":
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f
        else:
            assert np.array_equal(a, b), f


def _run_both(dev, oracle, model_path, elites_path, lam, keep, seed, threads=4, sigma=0.12, jump_p=0.02,
              imm_p=0.01, model_w=None):
    if model_w is None:
        m = scr.load_model(model_path)
        mo = oracle.load_model(model_path)
    else:
        m = scr.Model.from_compact(model_w)
        mo = oracle.model_from_compact(model_w)
    el = scr.load_elites(elites_path)
    elo = oracle.load_elites(elites_path)
    cfg = scr.Cfg(lam, keep, seed, threads, float(np.float32(sigma)), float(np.float32(jump_p)),
                  float(np.float32(imm_p)))
    print(f"case mode={dev.mode} lam={lam} keep={keep} seed={seed} T={threads} sigma={sigma} jump={jump_p} imm={imm_p} "
          f"ne={len(el)}", flush=True)
    got, screens = dev.run(m, el, cfg)
    ref, rscreens = oracle.screen(mo, elo, lam, keep, seed, threads, cfg.sigma, cfg.jump_p, cfg.imm_p)
    assert screens == rscreens == lam
    _same(got, ref)
    return got, dev.stats()


@pytest.mark.parametrize("threads", [4, 1])
def test_config0_parity(dev, oracle, golden, threads):
    m, e = _c0(golden)
    got, st = _run_both(dev, oracle, m, e, 1 << 20, 256, 0, threads, 0.12, 0.01, 0.01)
    assert st["candidates"] == 1 << 20
    assert st["offset_path"] == (1 if dev.mode == "auto" else 0)
    assert np.all(np.diff(got["score"]) <= 0)


def test_reference_fixtures_parity(dev, oracle, golden):
    # Makefile `make screen` inputs (161-term model, 96 v6 elites) at 1e6
    _run_both(dev, oracle, os.path.join(golden, "surrogate_161.txt"), os.path.join(golden, "seed_pop.txt"),
              1_000_000, 192, 7, 4)


@pytest.mark.parametrize("lam,keep,threads", [(1, 5, 4), (3, 5, 4), (7, 3, 8), (1000, 2000, 4), (5000, 1, 3),
                                              (65536, 64, 64), (300000, 100, 7)])
def test_sizes_and_partitions(dev, oracle, golden, lam, keep, threads):
    m, e = _c0(golden)
    _run_both(dev, oracle, m, e, lam, keep, 12345, threads)


@pytest.mark.parametrize("sigma,jump_p,imm_p", [(0.0, 0.0, 0.0), (0.3, 0.0, 1.0), (0.05, 1.0, 0.0),
                                                (0.5, 0.5, 0.5), (0.12, 0.02, 0.01)])
def test_rates(dev, oracle, golden, sigma, jump_p, imm_p):
    m, e = _c0(golden)
    _run_both(dev, oracle, m, e, 200000, 128, 99, 4, sigma, jump_p, imm_p)


@pytest.mark.parametrize("n", [153, 159, 161, 164, 165, 166])
def test_model_widths(dev, oracle, golden, n):
    w = (np.random.default_rng(n).standard_normal(n) / np.sqrt(n)).astype(np.float32)
    _, e = _c0(golden)
    _run_both(dev, oracle, None, e, 100000, 64, n, 4, model_w=w)


def test_all_ties(dev, oracle, golden):
    # zero model: every score equal; survivors are the first candidates of
    # each stream in heap order, t-major (exact heap replay)
    _, e = _c0(golden)
    got, _ = _run_both(dev, oracle, None, e, 1 << 20, 300, 5, 4, model_w=np.zeros(153, np.float32))
    assert list(got["stream"][:256]) == [0] * 256
    # constant intercept only: same ties, different value
    w = np.zeros(153, np.float32)
    w[0] = 0.75
    _run_both(dev, oracle, None, e, 100000, 300, 6, 4, model_w=w)


@pytest.mark.parametrize("elites", ["lds", "global"])
def test_signed_zero_model_scores(dev, oracle, golden, tmp_path, elites):
    # every weight -0.0: every product is a signed zero.  The reference's
    # sums start at +0, so every score is +0; the scorer starts its sums at
    # the first product and must still return +0 (with fewer than 48 KB of
    # elites they sit in LDS, with 700 the scorer reads them from global
    # memory and the model through scalar loads)
    _, e = _c0(golden)
    if elites == "global":
        rng = np.random.default_rng(12)
        pop = tmp_path / "big.txt"
        with open(pop, "w") as f:
            for i in range(700):
                genes = " ".join(f"{x:.4f}" for x in rng.random(28))
                f.write(f"{i} 1 2 3 {i % 12} {genes}\n")
        e = str(pop)
    # width 166: the linear tail over genes 16..28 is -0.0 too (at width 153
    # it is +0, which alone would restore the sign of a zero score)
    w = np.full(166, -0.0, np.float32)
    got, _ = _run_both(dev, oracle, None, e, 1 << 18, 200, 9, 4, model_w=w)
    assert not np.any(np.signbit(got["score"]))
    # intercept -0.0 alone, the other weights nonzero
    w = np.random.default_rng(13).standard_normal(153).astype(np.float32)
    w[0] = -0.0
    _run_both(dev, oracle, None, e, 1 << 18, 200, 10, 4, model_w=w)


def test_coarse_scores_many_ties(dev, oracle, golden, tmp_path):
    # weights on a coarse lattice and elites on a coarse lattice with no
    # mutation: many exactly equal scores straddle the cut
    rng = np.random.default_rng(8)
    w = rng.integers(-2, 3, 153).astype(np.float32)
    pop = tmp_path / "coarse.txt"
    with open(pop, "w") as f:
        for i in range(8):
            genes = " ".join(str(x) for x in rng.integers(0, 2, 16))
            f.write(f"{i} 0 0 0 {i} {genes}\n")
    _run_both(dev, oracle, None, str(pop), 400000, 500, 21, 4, sigma=0.0, jump_p=0.0, imm_p=0.0, model_w=w)


def test_many_elites_global_path(dev, oracle, golden, tmp_path):
    rng = np.random.default_rng(2)
    pop = tmp_path / "big.txt"
    with open(pop, "w") as f:
        for i in range(700):
            genes = " ".join(f"{x:.4f}" for x in rng.random(28))
            f.write(f"{i} 1 2 3 {i % 12} {genes}\n")  # 33-token v5 lines
    m, _ = _c0(golden)
    _run_both(dev, oracle, m, str(pop), 300000, 150, 4, 4)


def test_screen_cmd_files_match_oracle(oracle, golden, tmp_path):
    m, e = _c0(golden)
    out, cnt = str(tmp_path / "promoted.txt"), str(tmp_path / "screens.txt")
    rc = scr.screen_cmd(["--model", m, "--elites", e, "--out", out, "--count-out", cnt, "--lambda", "300000",
                         "--keep", "192", "--seed", "77", "--threads", "4", "--sigma", "0.1200", "--jump", "0.02",
                         "--imm", "0.01"])
    assert rc == 0
    ref, total = oracle.screen(oracle.load_model(m), oracle.load_elites(e), 300000, 192, 77, 4,
                               float(np.float32(0.12)), float(np.float32(0.02)), float(np.float32(0.01)))
    oracle.write_promoted(str(tmp_path / "ref.txt"), ref)
    assert open(out, "rb").read() == open(tmp_path / "ref.txt", "rb").read()
    assert open(cnt).read() == f"{total}\n" == "300000\n"


@pytest.mark.parametrize("threads", [4, 1])
def test_record_overflow_split_and_retry(dev, oracle, golden, threads):
    # Record buffers of 2048 records (one block emits at most 1987): the first
    # launches emit every candidate of 1 and 3 blocks per stream while the
    # heaps fill, so they overflow and are split -- by streams while more than
    # one stream is active, then by blocks inside one stream -- and redone
    # with the other record buffer's launch still queued behind them.
    m, e = _c0(golden)
    dev.set_record_cap(2048)
    try:
        _, st = _run_both(dev, oracle, m, e, 1 << 20, 256, 41, threads, 0.12, 0.01, 0.01)
    finally:
        dev.set_record_cap(1 << 20)
    assert st["retries"] >= (6 if threads == 4 else 2), st
    with pytest.raises(CsError):
        dev.set_record_cap(1024)


def test_keep_zero_and_empty(dev, golden):
    m, e = _c0(golden)
    with pytest.raises(CsError):
        dev.run(scr.load_model(m), scr.load_elites(e), scr.Cfg(1000, 0, 1))
    got, screens = dev.run(scr.load_model(m), scr.load_elites(e), scr.Cfg(0, 10, 1))
    assert len(got) == 0 and screens == 0


def test_allocation_failure_is_out_of_memory_and_the_context_recovers(dev, oracle, golden):
    # 2^60 candidates need about 9 PB of block entries: the device allocation
    # fails before any launch and the screen reports OutOfMemory; the same
    # context then screens normally (the failed allocation leaves no error
    # state behind for the next call's launch checks).
    m, e = _c0(golden)
    cfg = scr.Cfg(1 << 60, 16, 0, 4, float(np.float32(0.12)), float(np.float32(0.01)), float(np.float32(0.01)))
    print(f"case mode={dev.mode} lam=2^60 (allocation must fail)", flush=True)
    with pytest.raises(CsError) as ei:
        dev.run(scr.load_model(m), scr.load_elites(e), cfg)
    assert ei.value.name == "OutOfMemory", str(ei.value)
    _run_both(dev, oracle, m, e, 1 << 16, 64, 3, 4, 0.12, 0.01, 0.01)


def test_large_lambda_parity(dev, oracle, golden):
    # 2^26 candidates in 4 streams: every survivor bit-exact against the oracle
    m, e = _c0(golden, "c1_elites.txt")
    _run_both(dev, oracle, m, e, 1 << 26, 128, 3, 4, 0.12, 0.01, 0.01)


def test_production_generation_properties(dev, oracle, golden):
    # one full 2.65e9-candidate generation (128 elites, keep 128, T=4):
    # count, order, and every survivor re-synthesized by the oracle from its
    # (stream, start) and re-scored bit for bit
    m, e = _c0(golden, "c1_elites.txt")
    model, el = scr.load_model(m), scr.load_elites(e)
    mo, elo = oracle.load_model(m), oracle.load_elites(e)
    lam = 2_650_000_000
    cfg = scr.Cfg(lam, 128, 0, 4, float(np.float32(0.12)), float(np.float32(0.01)), float(np.float32(0.01)))
    print("case production generation lam=2.65e9", flush=True)
    got, screens = dev.run(model, el, cfg)
    st = dev.stats()
    assert screens == lam and st["candidates"] == lam
    assert st["offset_path"] == (1 if dev.mode == "auto" else 0)
    assert len(got) == 128 and np.all(np.diff(got["score"]) <= 0)
    per = lam // 4
    for s in got:
        assert s["index"] < (per if s["stream"] < 3 else lam - 3 * per)
        g, L, ia, ib = oracle.synth(elo, oracle.stream_seed(0, int(s["stream"])), int(s["start"]), cfg.sigma,
                                    cfg.jump_p, cfg.imm_p)
        assert np.array_equal(g.view(np.uint32), s["g"].view(np.uint32))
        assert np.float32(oracle.score(mo, g)).view(np.uint32) == s["score"].view(np.uint32)
        assert (elo[ia].id, elo[ib].id, elo[ia].tribe) == (s["pa"], s["pb"], s["tribe"])
    print(f"stats {st}", flush=True)


def test_offset_paths_agree_at_production_size(golden):
    # the speculative walk (verified) and the transfer maps are two
    # independent ways to place every candidate; a full generation must give
    # the same survivors either way
    m, e = _c0(golden, "c1_elites.txt")
    model, el = scr.load_model(m), scr.load_elites(e)
    cfg = scr.Cfg(2_650_000_000, 128, 1, 4, float(np.float32(0.1116)), float(np.float32(0.01)),
                  float(np.float32(0.01)))
    out = {}
    for mode in (0, 1):
        print(f"case production generation offset mode {mode}", flush=True)
        with scr.Screen(0) as d:
            d.set_offset_mode(mode)
            out[mode], _ = d.run(model, el, cfg)
            assert d.stats()["offset_path"] == (1 - mode)
    _same(out[0], out[1])


def test_slow_mutation_path_edge_inputs(dev, oracle, golden, tmp_path):
    # a subnormal sigma (sigma * 2^-24 not representable) and a non-finite
    # parent gene (clamp01(NaN) = 1) take the general mutation path
    m, e = _c0(golden)
    _run_both(dev, oracle, m, e, 100000, 64, 31, 4, sigma=1e-40, jump_p=0.02, imm_p=0.01)
    lines = open(e).read().splitlines()
    toks = lines[0].split()
    toks[7] = "nan"
    toks[9] = "inf"
    pop = tmp_path / "nan.txt"
    pop.write_text("\n".join([" ".join(toks)] + lines[1:]) + "\n")
    got, _ = _run_both(dev, oracle, m, str(pop), 100000, 64, 32, 4)
    assert np.all((got["g"] >= 0) & (got["g"] <= 1))


def test_arena_executable_as_evolve_calls_it(oracle, golden, tmp_path):
    # brain/evolve.py:421-438 restated: the elites file it writes (genome_line,
    # 34-token v6 lines of the top E), then `<zig> screen ...` as a subprocess
    # with exactly its argv; both output files must equal the oracle writer's.
    import subprocess
    import sys
    pop, seed, shard, gen, screens, threads = 96, 7, 3, 2, 300_000, 4
    E = max(2, pop // 8)
    K = pop - E - max(2, pop // 8)
    lines = open(os.path.join(golden, "seed_pop.txt")).read().splitlines()[:E]
    elites = tmp_path / "elites.txt"
    elites.write_text("".join(line + "\n" for line in lines))
    model = os.path.join(golden, "surrogate_161.txt")
    prom, cnt = tmp_path / "promoted.txt", tmp_path / "screens.txt"
    sigma = f"{0.12 * (0.93 ** gen):.4f}"
    argv = ["screen", "--model", model, "--elites", str(elites), "--out", str(prom), "--count-out", str(cnt),
            "--lambda", str(screens), "--keep", str(K), "--seed", str(seed * 1000003 + shard * 131 + gen),
            "--threads", str(threads), "--sigma", sigma, "--jump", "0.02", "--imm", "0.01"]
    exe = os.path.join(os.path.dirname(golden), "..", "codec-arena_amd", "bin", "arena")
    print(f"case arena subprocess K={K} sigma={sigma}", flush=True)
    env = dict(os.environ)
    env.pop("CSFUNNEL_PROMOTED_FORMAT", None)
    r = subprocess.run([sys.executable, exe] + argv, capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stderr
    assert r.stderr.strip().endswith(f"screen: {screens} candidates scored, top {K} promoted")
    cfg = scr.caller_cfg(screens, pop, seed, shard, gen, threads)
    assert (cfg.keep, cfg.seed) == (K, seed * 1000003 + shard * 131 + gen)
    ref, total = oracle.screen(oracle.load_model(model), oracle.load_elites(str(elites)), cfg.lambda_, cfg.keep,
                               cfg.seed, threads, cfg.sigma, cfg.jump_p, cfg.imm_p)
    oracle.write_promoted(str(tmp_path / "ref.txt"), ref)
    assert prom.read_bytes() == (tmp_path / "ref.txt").read_bytes()
    assert cnt.read_text() == f"{total}\n" == f"{screens}\n"
    # the named v6 choice: evolve.py keeps these lines (all 29 genes)
    env["CSFUNNEL_PROMOTED_FORMAT"] = "v6"
    r = subprocess.run([sys.executable, exe] + argv + ["--v6-out", str(tmp_path / "v6b.txt")], capture_output=True,
                       text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stderr
    v6 = prom.read_text().splitlines()
    assert v6 == (tmp_path / "v6b.txt").read_text().splitlines()
    assert len(v6) == K and all(len(line.split()) == 34 for line in v6)
    for line, zline, s in zip(v6, open(tmp_path / "ref.txt"), ref):
        assert line.split()[:29] == zline.split()
        assert [float(x) for x in line.split()[5:]] == [float(oracle.fmt(float(x))) for x in s["g"]]


def _chain_elites(survivors, gen):
    """Survivors of generation gen as the next generation's elites (ids
    128 gen + rank, genes 0..22 and regulators from the survivor's 29 genes),
    for the device and for the oracle."""
    from csfunnel import _abi
    from oracle_binding import OrGenome
    n = len(survivors)
    dev_el, or_el = (_abi.Genome * n)(), (OrGenome * n)()
    for r, s in enumerate(survivors):
        for e in (dev_el[r], or_el[r]):
            e.id, e.gen = gen * 128 + r, gen + 1
            e.pa, e.pb, e.tribe = int(s["pa"]), int(s["pb"]), int(s["tribe"])
            for k in range(23):
                e.g[k] = float(s["g"][k])
            for k in range(6):
                e.reg[k] = float(s["g"][23 + k])
    return dev_el, or_el


@pytest.mark.timeout(1200)
@pytest.mark.parametrize("last", [1, 27])
def test_production_chained_generation_bit_exact(oracle, golden, last):
    # The production shard at full size and over its chain: generations
    # 0 .. last - 1 (2.65e9 candidates each, 128 elites, keep 128, 4 streams,
    # seed g, sigma "%.4f" % (0.12 * 0.93^g)) on the GPU, each generation's
    # survivors the next one's elites; generation `last` on the GPU must then
    # equal the oracle's -- every survivor field and screens_count -- over all
    # 2.65e9 candidates.  last = 27 is the 28th and final generation of the
    # shard's run, where sigma has annealed to 0.0171 and the elites are
    # 27 generations of GPU survivors.
    m, e = _c0(golden, "c1_elites.txt")
    model, mo = scr.load_model(m), oracle.load_model(m)
    lam = 2_650_000_000
    el, or_el = scr.load_elites(e), None
    with scr.Screen(0) as d:
        for g in range(last + 1):
            sigma = scr.parse_f32(f"{0.12 * 0.93 ** g:.4f}")
            cfg = scr.Cfg(lam, 128, g, 4, sigma, float(np.float32(0.01)), float(np.float32(0.01)))
            print(f"case chained production generation {g} (GPU)", flush=True)
            surv, screens = d.run(model, el, cfg)
            assert d.stats()["candidates"] == lam and screens == lam
            if g < last:
                el, or_el = _chain_elites(surv, g)
    got = surv
    print(f"case chained production generation {last} (oracle, 4 streams in 64 chunks each on 16 host threads)",
          flush=True)
    ref, rscreens = oracle.screen_chunked(mo, or_el, lam, 128, last, 4, cfg.sigma, cfg.jump_p, cfg.imm_p,
                                          workers=16, chunks=64)
    assert rscreens == lam
    _same(got, ref)


@pytest.mark.parametrize("sigma", [0.0, 0.12, 1e-40])
def test_negative_zero_parent_genes(dev, oracle, golden, tmp_path, sigma):
    # parents read as -0.0: with sigma 0 half of the mutated genes are
    # -0 + -0 = -0 before the clamp, which gives +0 (clamp01(-0) = +0)
    rng = np.random.default_rng(3)
    pop = tmp_path / "negzero.txt"
    with open(pop, "w") as f:
        for i in range(16):
            genes = ["-0.0000" if rng.random() < 0.5 else f"{x:.4f}" for x in rng.random(16)]
            f.write(f"{i} 0 0 0 {i % 12} " + " ".join(genes) + "\n")
    m, _ = _c0(golden)
    got, _ = _run_both(dev, oracle, m, str(pop), 200000, 128, 8, 4, sigma=sigma, jump_p=0.0, imm_p=0.0)
    assert not np.any(np.signbit(got["g"]))
