"""GPU FP64 ridge refit and the ridge-loss backward kernel against the numpy
oracle (oracle/ridge_oracle.py) at config-2 size: 356,377 genomes, 153
features, alpha = 1e-3; plus the brain's 166-term refit toward a prior."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import ridge_oracle as ro  # noqa: E402

pytestmark = pytest.mark.gpu

N_C2 = 356_377


@pytest.fixture(scope="module")
def c2():
    g = np.random.default_rng(0).random((N_C2, 16))
    X = ro.features(g)
    c = np.random.default_rng(1).standard_normal(153) / np.sqrt(153.0)
    y = X @ c + np.random.default_rng(2).normal(0.0, 0.5, N_C2)
    return g, X, y, c


@pytest.fixture(scope="module")
def dev():
    from csfunnel import screen as scr
    d = scr.Screen(0)
    yield d
    d.close()


def test_gram_matches_numpy(c2, dev):
    import torch
    from csfunnel import ridge
    g, X, y, _ = c2
    print("case gram n=356377 p=153", flush=True)
    G, xty, yy = ridge.gram(torch.from_numpy(g), torch.from_numpy(y), dev)
    G, xty = G.cpu().numpy(), xty.cpu().numpy()
    Gr, br = X.T @ X, X.T @ y
    assert np.max(np.abs(G - Gr)) <= 1e-12 * np.max(np.abs(Gr))
    assert np.array_equal(G, G.T)
    assert np.max(np.abs(xty - br)) <= 1e-12 * np.max(np.abs(br))
    assert abs(float(yy) - y @ y) <= 1e-12 * (y @ y)


def test_refit_matches_numpy_solve(c2, dev):
    import torch
    from csfunnel import ridge
    g, X, y, c = c2
    print("case refit alpha=1e-3", flush=True)
    w = ridge.refit(torch.from_numpy(g), torch.from_numpy(y), 1e-3, None, dev).cpu().numpy()
    wr = ro.refit(X, y, 1e-3)
    err = np.abs(w - wr)
    print(f"per-coefficient |dw|: max {err.max():.3e} median {np.median(err):.3e} "
          f"(|w| max {np.abs(wr).max():.3f}); recovery |w - c| max {np.abs(wr - c).max():.3f}", flush=True)
    cond = np.linalg.cond(X.T @ X + 1e-3 * np.eye(153))
    assert err.max() <= 1e-15 * cond * (1 + np.abs(wr).max())


def test_loss_grad_kernel_matches_numpy_and_fd(c2, dev):
    import torch
    from csfunnel import ridge
    g, X, y, _ = c2
    rng = np.random.default_rng(7)
    w = rng.standard_normal(153) * 0.1
    prior = rng.standard_normal(153) * 0.1
    tg, ty = torch.from_numpy(g).cuda(), torch.from_numpy(y).cuda()
    tp = torch.from_numpy(prior).cuda()
    print("case loss/grad", flush=True)
    L, G = ridge.loss_and_grad(torch.from_numpy(w), tg, ty, 1e-3, tp, dev)
    Lr, Gr = ro.loss_grad(X, y, w, 1e-3, prior)
    assert abs(float(L) - Lr) <= 1e-12 * Lr
    assert np.max(np.abs(G.cpu().numpy() - Gr)) <= 1e-11 * np.max(np.abs(Gr))
    # central finite differences of the GPU loss, h = 1e-4 (FP64)
    h = 1e-4
    for k in (0, 3, 16, 17, 100, 152):
        e = np.zeros(153)
        e[k] = h
        Lp, _ = ridge.loss_and_grad(torch.from_numpy(w + e), tg, ty, 1e-3, tp, dev)
        Lm, _ = ridge.loss_and_grad(torch.from_numpy(w - e), tg, ty, 1e-3, tp, dev)
        fd = (float(Lp) - float(Lm)) / (2 * h)
        assert abs(fd - float(G[k])) <= 1e-6 * (1 + abs(float(G[k]))), (k, fd, float(G[k]))


def test_gradient_vanishes_at_gpu_refit_and_autograd(c2, dev):
    import torch
    from csfunnel import ridge
    g, X, y, _ = c2
    tg, ty = torch.from_numpy(g).cuda(), torch.from_numpy(y).cuda()
    w_star = ridge.refit(tg, ty, 1e-3, None, dev)
    _, g0 = ridge.loss_and_grad(torch.zeros(153, dtype=torch.float64), tg, ty, 1e-3, None, dev)
    _, gs = ridge.loss_and_grad(w_star, tg, ty, 1e-3, None, dev)
    assert float(gs.abs().max()) <= 1e-8 * float(g0.abs().max())
    # autograd.Function: backward returns the kernel's gradient
    w = torch.randn(153, dtype=torch.float64, device="cuda", requires_grad=True)
    loss = ridge.ridge_loss(w, tg, ty, 1e-3, None, dev)
    loss.backward()
    _, gk = ridge.loss_and_grad(w.detach(), tg, ty, 1e-3, None, dev)
    assert torch.equal(w.grad, gk)
    # a few steps of plain gradient descent through autograd reduce the loss
    G, _, _ = ridge.gram(tg, ty, dev)
    lr = 1.0 / float(torch.linalg.eigvalsh(2.0 * G / tg.shape[0]).max())  # stable step for this quadratic
    w2 = torch.zeros(153, dtype=torch.float64, device="cuda", requires_grad=True)
    opt = torch.optim.SGD([w2], lr=lr)
    losses = []
    for _ in range(20):
        opt.zero_grad()
        l2 = ridge.ridge_loss(w2, tg, ty, 1e-3, None, dev)
        l2.backward()
        opt.step()
        losses.append(float(l2.detach()))
    assert all(b <= a for a, b in zip(losses, losses[1:])) and losses[-1] < 0.9 * losses[0]


def test_refit_surrogate_166_toward_prior(dev, tmp_path):
    from csfunnel import ridge
    rng = np.random.default_rng(11)
    genes = rng.random((4000, 29))
    fit = rng.normal(30, 2, 4000)
    prior = list(rng.standard_normal(166))
    out = str(tmp_path / "surrogate_live.txt")
    rows = [(list(a), float(b)) for a, b in zip(genes, fit)]
    assert not ridge.refit_surrogate(rows[:299], prior, out, 50.0, dev)
    assert ridge.refit_surrogate(rows, prior, out, 50.0, dev)
    w = np.array(ridge.load_weights(out))
    wr = ro.refit(ro.features(genes), fit, 50.0, np.array(prior))
    assert np.max(np.abs(w - wr)) <= 1e-6 * (1 + np.abs(wr).max())  # %.8e text
    assert len(open(out).read().split()) == 167


def test_not_positive_definite_is_an_error(dev):
    import torch
    from csfunnel import ridge
    from csfunnel._abi import CsError
    g = torch.rand(20, 16, dtype=torch.float64)
    y = torch.rand(20, dtype=torch.float64)
    G, xty, _ = ridge.gram(g, y, dev)
    with pytest.raises(CsError):
        ridge.solve(G, xty, 0.0, None, dev)
