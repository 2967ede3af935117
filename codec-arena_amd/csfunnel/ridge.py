"""FP64 surrogate refit and the differentiable ridge loss on the GPU.

Host mirror of the brain's refit (codec-arena snapshot):
  load_weights / save_weights  brain/evolve.py:223-247
  refit_surrogate              brain/evolve.py:250-268
The work runs in libcsfunnel (cs_ridge_gram_f64 -> FP64 MFMA Gram assembly,
cs_ridge_solve_f64 -> Cholesky, cs_ridge_loss_grad_f64 -> loss and gradient);
torch supplies device memory.
"""
from __future__ import annotations

from typing import Optional, Sequence

import torch

from ._abi import check, lib
from .screen import Screen

_default: Optional[Screen] = None


def _dev(dev: Optional[Screen]) -> Screen:
    global _default
    if dev is not None:
        return dev
    if _default is None:
        _default = Screen(0)
    return _default


def _ng(genes: torch.Tensor) -> int:
    if genes.dim() != 2 or genes.shape[1] not in (16, 29):
        raise ValueError("genes must be (n, 16) or (n, 29)")
    return int(genes.shape[1])


def _f64(t) -> torch.Tensor:
    t = torch.as_tensor(t, dtype=torch.float64, device="cuda")
    return t.contiguous()


def gram(genes, y, dev: Optional[Screen] = None):
    """X'X, X'y, y'y of the features of `genes` (brain/evolve.py:264-265)."""
    d = _dev(dev)
    genes, y = _f64(genes), _f64(y)
    ng = _ng(genes)
    p = 1 + ng + 136
    G = torch.empty((p, p), dtype=torch.float64, device="cuda")
    xty = torch.empty(p, dtype=torch.float64, device="cuda")
    yy = torch.empty(1, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    check(lib().cs_ridge_gram_f64(d._h, genes.data_ptr(), ng, y.data_ptr(), genes.shape[0], G.data_ptr(),
                                  xty.data_ptr(), yy.data_ptr()))
    return G, xty, yy[0]


def solve(G, xty, alpha: float, prior=None, dev: Optional[Screen] = None) -> torch.Tensor:
    """solve(G + alpha I, xty + alpha prior) (brain/evolve.py:266)."""
    d = _dev(dev)
    G, xty = _f64(G), _f64(xty)
    p = xty.shape[0]
    w = torch.empty(p, dtype=torch.float64, device="cuda")
    pr = _f64(prior) if prior is not None else None
    torch.cuda.synchronize()
    check(lib().cs_ridge_solve_f64(d._h, G.data_ptr(), xty.data_ptr(), p, float(alpha),
                                   pr.data_ptr() if pr is not None else None, w.data_ptr()))
    return w


def refit(genes, y, alpha: float, prior=None, dev: Optional[Screen] = None) -> torch.Tensor:
    G, xty, _ = gram(genes, y, dev)
    return solve(G, xty, alpha, prior, dev)


def loss_and_grad(w, genes, y, alpha: float, prior=None, dev: Optional[Screen] = None):
    d = _dev(dev)
    w, genes, y = _f64(w), _f64(genes), _f64(y)
    ng = _ng(genes)
    loss = torch.empty(1, dtype=torch.float64, device="cuda")
    grad = torch.empty_like(w)
    pr = _f64(prior) if prior is not None else None
    torch.cuda.synchronize()
    check(lib().cs_ridge_loss_grad_f64(d._h, genes.data_ptr(), ng, y.data_ptr(), genes.shape[0], w.data_ptr(),
                                       pr.data_ptr() if pr is not None else None, float(alpha), loss.data_ptr(),
                                       grad.data_ptr()))
    return loss[0], grad


class RidgeLoss(torch.autograd.Function):
    """L(w) = (|X w - y|^2 + alpha |w - prior|^2) / n with the hand-written
    gradient kernel as its backward; minimized by refit()."""

    @staticmethod
    def forward(ctx, w, genes, y, alpha, prior=None, dev=None):
        loss, grad = loss_and_grad(w.detach(), genes, y, alpha, prior, dev)
        ctx.save_for_backward(grad)
        return loss

    @staticmethod
    def backward(ctx, gout):
        (grad,) = ctx.saved_tensors
        return gout * grad, None, None, None, None, None


def ridge_loss(w, genes, y, alpha, prior=None, dev=None):
    return RidgeLoss.apply(w, genes, y, alpha, prior, dev)


def load_weights(path: str) -> list:
    """brain/evolve.py:223-241: read a model file, migrating to 166 terms."""
    with open(path) as f:
        toks = f.read().split()
    n = int(toks[0])
    w = [float(x) for x in toks[1:1 + n]]
    inserts = {153: (17, 13), 159: (17, 7), 161: (19, 5), 164: (22, 2), 165: (29, 1)}
    if n in inserts:
        at, k = inserts[n]
        w = w[:at] + [0.0] * k + w[at:]
    return w


def save_weights(path: str, w: Sequence[float]) -> None:
    """brain/evolve.py:244-247: "<n>\\n" + "%.8e" values."""
    with open(path, "w") as f:
        f.write(f"{len(w)}\n")
        f.write(" ".join(f"{float(x):.8e}" for x in w) + "\n")


def refit_surrogate(rows, prior_w, out_path: str, alpha: float = 50.0, dev: Optional[Screen] = None) -> bool:
    """brain/evolve.py:250-268 on the GPU: rows = [(29 genes, fitness)];
    needs >= 300 rows and a 166-term prior; writes the new model file."""
    if len(rows) < 300:
        return False
    genes = torch.tensor([list(g) for g, _ in rows], dtype=torch.float64)
    if genes.shape[1] != 29 or len(prior_w) != 166:
        return False
    y = torch.tensor([f for _, f in rows], dtype=torch.float64)
    w = refit(genes, y, alpha, torch.tensor(list(prior_w), dtype=torch.float64), dev)
    save_weights(out_path, w.cpu().tolist())
    return True
