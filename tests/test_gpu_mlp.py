"""Learned surrogate on the GPU: BF16 MFMA forward/backward against torch
references, AdamW against torch.optim.AdamW, training gain over the quadratic
fallback, and the BF16/FP8 MLP screens against FP64 scoring.

Training pairs are the codec label set (codec-arena_amd/data): PSNR - 14 bpp
of each genome from the codec label source on the ten-clip corpus."""
import os

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    from csfunnel import screen as scr
    d = scr.Screen(0)
    yield d
    d.close()


def bf16(x):
    import torch
    return torch.as_tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float32)


def torch_forward(p, X, dtype, round_acts):
    """Reference forward: BF16-rounded weights and inputs; activations rounded
    to BF16 between layers when round_acts (as the training kernels store them)."""
    import torch
    import torch.nn.functional as F
    from csfunnel.mlp import unpack
    W1, b1, W2, b2, w3, b3 = [torch.as_tensor(np.array(t), dtype=torch.float32) for t in unpack(p)]
    W1, W2 = bf16(W1), bf16(W2)
    x = bf16(X)
    cast = (lambda t: bf16(t)) if round_acts else (lambda t: t)
    h1 = cast(F.gelu((x.to(dtype) @ W1.to(dtype)) + b1.to(dtype), approximate="tanh"))
    h2 = cast(F.gelu(h1.to(dtype) @ W2.to(dtype) + b2.to(dtype), approximate="tanh"))
    return h2.to(dtype) @ w3.to(dtype) + float(b3)


def test_forward_matches_torch(dev):
    import torch
    from csfunnel.mlp import MLP
    m = MLP(dev, seed=3)
    X = torch.rand(5000, 29)
    out = m.forward(X.cuda()).cpu()
    ref = torch_forward(m.params, X[:, :16], torch.float32, True)
    err = (out - ref).abs().max().item()
    print(f"forward max |gpu - torch| = {err:.3e} (|out| max {ref.abs().max().item():.3f})", flush=True)
    assert err <= 2e-3 * (1 + ref.abs().max().item())


def test_gradients_match_fp64_autograd_and_fd(dev):
    import torch
    import torch.nn.functional as F
    from csfunnel.mlp import MLP, unpack
    m = MLP(dev, seed=4)
    X = torch.rand(2048, 16)
    y = torch.randn(2048)
    loss, g = m.grad(X.cuda(), y.cuda())
    # fp64 autograd on the same (bf16-rounded) weights and inputs
    parts = [torch.tensor(np.array(t), dtype=torch.float64) for t in unpack(m.params)]
    parts[0] = bf16(parts[0]).double()
    parts[2] = bf16(parts[2]).double()
    ts = [t.clone().requires_grad_(True) for t in parts[:5]]
    b3 = torch.tensor(float(parts[5]), dtype=torch.float64, requires_grad=True)
    x = bf16(X).double()

    def f(W1, b1, W2, b2, w3, b3):
        h1 = F.gelu(x @ W1 + b1, approximate="tanh")
        h2 = F.gelu(h1 @ W2 + b2, approximate="tanh")
        return ((h2 @ w3 + b3 - y.double()) ** 2).mean()

    L = f(*ts, b3)
    L.backward()
    ref = torch.cat([t.grad.flatten() for t in ts] + [b3.grad.flatten()]).numpy()
    assert abs(loss - L.item()) <= 2e-2 * L.item()
    cos = float(g @ ref / (np.linalg.norm(g) * np.linalg.norm(ref)))
    print(f"loss gpu {loss:.6f} fp64 {L.item():.6f}; gradient cosine {cos:.6f}", flush=True)
    assert cos > 0.999
    off = 0
    for t in ts + [b3]:
        n = t.numel()
        a, b = g[off:off + n], ref[off:off + n]
        assert np.linalg.norm(a - b) <= 5e-2 * np.linalg.norm(b) + 1e-7, (off, n)
        off += n
    # central finite differences of the fp64 reference (h = 1e-4) confirm the
    # reference gradient the kernels were compared with
    with torch.no_grad():
        for (ti, flat) in ((2, 1234), (1, 17), (4, 99)):
            e = torch.zeros_like(ts[ti]).flatten()
            e[flat] = 1e-4
            e = e.reshape(ts[ti].shape)
            args_p = [t.detach() + (e if i == ti else 0) for i, t in enumerate(ts)]
            args_m = [t.detach() - (e if i == ti else 0) for i, t in enumerate(ts)]
            fd = (f(*args_p, b3.detach()) - f(*args_m, b3.detach())).item() / 2e-4
            assert abs(fd - ts[ti].grad.flatten()[flat].item()) <= 1e-6 * (1 + abs(fd))


def test_adamw_step_matches_torch(dev):
    import torch
    from csfunnel.mlp import MLP
    m = MLP(dev, seed=5)
    X, y = torch.rand(4096, 16).cuda(), torch.randn(4096).cuda()
    p0 = m.params
    _, g = m.grad(X, y)
    w = torch.nn.Parameter(torch.tensor(p0))
    opt = torch.optim.AdamW([w], lr=1e-3, weight_decay=0.01)
    w.grad = torch.tensor(g)
    opt.step()
    m.train_step(X, y, 1e-3, 0.01)
    assert np.max(np.abs(m.params - w.detach().numpy())) <= 1e-6


@pytest.fixture(scope="module")
def trained(dev):
    from csfunnel import make_labels, surrogate
    genes, y, ntr, _ = make_labels.load()
    print("case train 20000 codec-labelled pairs x 300 epochs", flush=True)
    m, hist = surrogate.train_mlp(dev, genes[:ntr], y[:ntr])
    w = surrogate.fit_quadratic(dev, genes[:ntr], y[:ntr])
    return m, w, genes[:ntr], y[:ntr], genes[ntr:], y[ntr:], hist


def test_training_beats_quadratic_fallback(dev, trained):
    """What the trained network gains over the method without it: Spearman
    rho and top-256 recall against 5,000 held-out codec labels, MLP vs the
    153-term quadratic refit (FP64 GPU ridge) on the same 20,000 pairs."""
    from csfunnel import surrogate
    m, w, gtr, ytr, gte, yte, hist = trained
    assert hist[-1] < 0.5 * hist[0]
    r = surrogate.holdout_report(m, w, gte, yte)
    print(f"held-out Spearman rho: MLP fp64 {r['spearman_mlp']:.4f} bf16 {r['spearman_mlp_bf16']:.4f} "
          f"quadratic {r['spearman_quad']:.4f}; top-256 recall: MLP {r['recall_mlp']:.3f} "
          f"bf16 {r['recall_mlp_bf16']:.3f} quadratic {r['recall_quad']:.3f}", flush=True)
    assert r["spearman_mlp"] > r["spearman_quad"] and r["spearman_mlp_bf16"] > r["spearman_quad"]
    assert r["recall_mlp"] >= r["recall_quad"]


def test_funnel_survivors_oracle_verified(dev, trained):
    """Config-4 funnel over 4 chained generations of 2.65e9 candidates: the
    quadratic screen, the BF16 MLP screen alone, and the learned funnel (the
    BF16 MLP screen's top 1024 re-ranked in FP64 by the mean of the MLP and
    quadratic predictions) breed from the same elites; the codec label source
    re-scores each screen's 64 survivors.  The learned funnel's survivors
    must score at least as well as the quadratic screen's in every
    generation, and every screen must beat the elites it was bred from."""
    from csfunnel import codec, screen as scr, surrogate
    m, w = trained[0], trained[1]
    el = scr.load_elites(os.path.join(REPO, "tests", "golden", "c1_elites.txt"))
    print("case funnel 4 generations x 2.65e9, oracle re-score of 3 x 64 survivors per generation", flush=True)
    res = surrogate.funnel_compare(dev, el, m, w, codec.Evaluator(), gens=4,
                                   progress=lambda r: print({k: round(v, 3) if isinstance(v, float) else v
                                                             for k, v in r.items()}, flush=True))
    print(f"elites mean {res['elites_mean']:.3f} max {res['elites_max']:.3f}; summary {res['summary']}", flush=True)
    for g in res["generations"]:
        for arm in ("quad", "mlp", "ensemble"):
            assert g[f"{arm}_mean"] > res["elites_mean"], (g["gen"], arm)
        assert g["ensemble_mean"] >= g["quad_mean"], g
    assert res["summary"]["ensemble"]["mean_of_means"] >= res["summary"]["quad"]["mean_of_means"]


def _screen_recall(dev, m, scorer, lam, keep):
    import torch
    from csfunnel import screen as scr
    from csfunnel.mlp import screen_ex, unpack
    g = os.path.join(REPO, "tests", "golden")
    el = scr.load_elites(os.path.join(g, "c1_elites.txt"))
    cfg = scr.Cfg(lam, keep, 5, 4, float(np.float32(0.12)), float(np.float32(0.01)), float(np.float32(0.01)))
    surv, n, dg, ds = screen_ex(dev, None, el, cfg, scorer, m, dump=True)
    assert n == lam and len(surv) == keep
    # survivors are the top-`keep` dumped scores, in order
    order = torch.argsort(ds, descending=True)[:keep].cpu().numpy()
    per = lam // 4
    got = surv["stream"].astype(np.int64) * per + surv["index"].astype(np.int64)
    assert set(got.tolist()) == set(order.tolist())
    assert np.array_equal(surv["score"], ds.cpu().numpy()[got])
    # FP64 scoring of the identical genomes
    W1, b1, W2, b2, w3, b3 = [torch.tensor(np.array(t), dtype=torch.float64, device="cuda") for t in unpack(m.params)]
    x = dg[:, :16].double()
    s64 = torch.nn.functional.gelu(torch.nn.functional.gelu(x @ W1 + b1, approximate="tanh") @ W2 + b2,
                                   approximate="tanh") @ w3 + b3
    top64 = set(torch.argsort(s64, descending=True)[:keep].cpu().numpy().tolist())
    recall = len(top64 & set(got.tolist())) / keep
    dev_abs = (ds.double() - s64).abs().max().item()
    return recall, dev_abs, dg, ds


def _funnel_recall(dev, m, scorer, lam, keep):
    import torch
    from csfunnel import screen as scr
    from csfunnel.mlp import forward_f64, screen_ex, screen_funnel
    g = os.path.join(REPO, "tests", "golden")
    el = scr.load_elites(os.path.join(g, "c1_elites.txt"))
    cfg = scr.Cfg(lam, keep, 5, 4, float(np.float32(0.12)), float(np.float32(0.01)), float(np.float32(0.01)))
    out, n = screen_funnel(dev, el, cfg, scorer, m, rerank=4)
    _, _, dg, _ = screen_ex(dev, None, el, cfg, scorer, m, dump=True)
    s64 = forward_f64(m.params, dg[:, :16].cpu().numpy())
    top = set(np.argsort(-s64, kind="stable")[:keep].tolist())
    per = lam // 4
    got = set((out["stream"].astype(np.int64) * per + out["index"].astype(np.int64)).tolist())
    return len(top & got) / keep


def test_mlp_screen_bf16_recall(dev, trained):
    m = trained[0]
    print("case MLP screen bf16 lam=2^22 keep=64", flush=True)
    recall, dmax, _, _ = _screen_recall(dev, m, 1, 1 << 22, 64)
    print(f"BF16 MLP screen: top-64 recall vs FP64 = {recall:.4f}, max |score - fp64| = {dmax:.3e}", flush=True)
    # Raw BF16 order can miss a near-tie at the cut: the FP64 scores of the
    # 64th and 65th candidates differ by ~1e-2 while BF16 rounding of the
    # genes alone moves scores by up to ~0.5 (CPU emulation of the rounding
    # steps), so at most one miss in 64 is allowed raw (63 of 64 measured,
    # max |score - fp64| 0.67); the funnel's FP64 re-rank of the BF16 top
    # 256 below must reach 0.99 (1.0 measured).
    assert recall >= 0.98
    frec = _funnel_recall(dev, m, 1, 1 << 22, 64)
    print(f"BF16 screen + FP64 re-rank of the top 256: top-64 recall vs FP64 = {frec:.4f}", flush=True)
    assert frec >= 0.99


def test_mlp_screen_fp8_recall(dev, trained):
    import torch
    m, gtr = trained[0], trained[2]
    m.calibrate_fp8(torch.tensor(gtr, dtype=torch.float32))
    print("case MLP screen fp8 lam=2^22 keep=64", flush=True)
    recall, dmax, _, _ = _screen_recall(dev, m, 2, 1 << 22, 64)
    print(f"FP8 MLP screen: top-64 recall vs FP64 = {recall:.4f}, max |score - fp64| = {dmax:.3e}", flush=True)
    # E4M3's 3-bit mantissa bounds the raw order (0.73 measured with W2 in
    # per-block E8M0 scales; per-tensor W2 scales gave 0.56); a wrong block
    # layout or scale puts scores off by powers of two and recall near 0
    assert recall >= 0.7
    assert dmax < 8.0
    frec = _funnel_recall(dev, m, 2, 1 << 22, 64)
    print(f"FP8 screen + FP64 re-rank of the top 256: top-64 recall vs FP64 = {frec:.4f}", flush=True)
    assert frec >= 0.95


def test_dump_matches_oracle_synthesis(dev, oracle):
    # every dumped genome of stream 0 and 3 equals the oracle's chain, and the
    # dumped quadratic scores are the oracle's, bit for bit
    from csfunnel import screen as scr
    from csfunnel.mlp import screen_ex
    g = os.path.join(REPO, "tests", "golden")
    mp, ep = os.path.join(g, "c0_model.txt"), os.path.join(g, "c0_elites.txt")
    lam = 1 << 20
    cfg = scr.Cfg(lam, 32, 9, 4, float(np.float32(0.12)), float(np.float32(0.02)), float(np.float32(0.01)))
    surv, n, dg, ds = screen_ex(dev, scr.load_model(mp), scr.load_elites(ep), cfg, 0, None, dump=True)
    dg, ds = dg.cpu().numpy(), ds.cpu().numpy()
    elo, mo = oracle.load_elites(ep), oracle.load_model(mp)
    per = lam // 4
    for t in (0, 3):
        seed = oracle.stream_seed(9, t)
        start = 0
        for i in range(3000):
            gg, L, _, _ = oracle.synth(elo, seed, start, cfg.sigma, cfg.jump_p, cfg.imm_p)
            row = t * per + i
            assert np.array_equal(gg.view(np.uint32), dg[row].view(np.uint32)), (t, i)
            assert np.float32(oracle.score(mo, gg)).view(np.uint32) == ds[row].view(np.uint32)
            start += L


def test_context_outlives_its_mlps():
    """cs_ctx_destroy refuses while an MLP handle is alive; Screen.close
    closes its MLPs first, so the MLP never syncs on a destroyed stream."""
    import ctypes as C
    from csfunnel import screen as scr
    from csfunnel._abi import lib
    from csfunnel.mlp import MLP
    d = scr.Screen(0)
    m = MLP(d, seed=1)
    assert lib().cs_ctx_destroy(d._h) == -3  # CS_E_BADARGS
    assert b"live MLP" in lib().cs_last_error()
    p = m.params
    d.close()
    assert m._h is None and d._h is None
    del p


@pytest.mark.parametrize("scorer", [1, 2])
def test_mlp_screen_survivor_genomes(dev, trained, scorer):
    """Genes 16..28 are synthesized only in batches that emit: the survivors of
    a plain screen (genomes included) equal those of the dumping screen, whose
    genes the oracle's synthesis confirms."""
    from csfunnel import screen as scr
    from csfunnel.mlp import screen_ex
    m = trained[0]
    g = os.path.join(REPO, "tests", "golden")
    el = scr.load_elites(os.path.join(g, "c1_elites.txt"))
    cfg = scr.Cfg(1 << 21, 256, 9, 4, float(np.float32(0.12)), float(np.float32(0.05)), float(np.float32(0.05)))
    print(f"case survivor genomes scorer={scorer} lam=2^21 keep=256", flush=True)
    a, n, dg, ds = screen_ex(dev, None, el, cfg, scorer, m, dump=True)
    b, n2, _, _ = screen_ex(dev, None, el, cfg, scorer, m)
    assert n == n2 == 1 << 21
    assert np.array_equal(a, b)
    per = (1 << 21) // 4
    rows = a["stream"].astype(np.int64) * per + a["index"].astype(np.int64)
    assert np.array_equal(a["g"], dg.cpu().numpy()[rows])


def test_f64_scorer_is_the_exact_fp64_top_k(dev, trained):
    """The FP64 scorer (FP64 MFMA layer 2) keeps the exact FP64 top-k of every
    candidate: its scores agree with the host FP64 reference within 1e-12,
    its survivors are the top-k of the host FP64 scores of every dumped
    genome, and the plain screen returns the same survivors as the dumping
    one."""
    from csfunnel import screen as scr
    from csfunnel.mlp import SCORER_MLP_F64, forward_f64, screen_ex, screen_f64
    m = trained[0]
    g = os.path.join(REPO, "tests", "golden")
    el = scr.load_elites(os.path.join(g, "c1_elites.txt"))
    lam, keep = 1 << 20, 64
    cfg = scr.Cfg(lam, keep, 5, 4, float(np.float32(0.12)), float(np.float32(0.01)), float(np.float32(0.01)))
    print("case FP64 scorer lam=2^20 keep=64", flush=True)
    keys = np.zeros(keep, np.float64)
    a, n, dg, ds = screen_ex(dev, None, el, cfg, SCORER_MLP_F64, m, dump=True, keys=keys)
    b, s64, n2 = screen_f64(dev, el, cfg, m)
    assert n == n2 == lam and len(a) == len(b) == keep
    assert np.array_equal(a, b) and np.array_equal(keys, s64)
    assert np.all(np.diff(s64) <= 0)
    ref = forward_f64(m.params, b["g"])
    err = np.abs(s64 - ref) / (1.0 + np.abs(ref))
    print(f"FP64 scorer: max relative |device - host| = {err.max():.3e}", flush=True)
    assert err.max() <= 1e-12
    assert np.array_equal(b["score"], s64.astype(np.float32))
    allref = forward_f64(m.params, dg[:, :16].cpu().numpy())
    order = np.argsort(-allref, kind="stable")
    per = lam // 4
    rows = b["stream"].astype(np.int64) * per + b["index"].astype(np.int64)
    assert np.array_equal(b["g"], dg.cpu().numpy()[rows])
    gap = allref[order[keep - 1]] - allref[order[keep]]
    print(f"host top-{keep} cut gap {gap:.3e}", flush=True)
    assert set(rows.tolist()) == set(order[:keep].tolist())
