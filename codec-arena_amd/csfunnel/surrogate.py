"""The learned-surrogate funnel: fit the MLP and its quadratic fallback on
codec-labelled genomes, judge both on held-out labels, and run both screens
on the GPU with their survivors re-scored by the codec label source.

The quadratic fallback is the reference's surrogate form: 153 terms over the
16 legacy effector genes (brain/evolve.py:211-220 `features`), refit by ridge
(brain/evolve.py:250-268; here FP64 on the GPU, ridge.refit) and screened
through Model.fromCompact (src/screen.zig:36-95).  The MLP has no counterpart
in the reference.
"""
from __future__ import annotations

import time
from typing import Callable, Dict, List, Optional

import numpy as np
import torch

from . import ridge
from .mlp import MLP, SCORER_MLP_BF16, forward_f64, screen_ex, screen_funnel, spearman, topk_recall
from .screen import Cfg, Model, Screen

N_QUAD = 153


def features16(genes) -> torch.Tensor:
    """[1, g_0..g_15, g_i g_j for i < 16 for j <= i] in FP64 (the 153-term
    layout of brain/evolve.py:211-220 restricted to the legacy genes)."""
    g = torch.as_tensor(genes, dtype=torch.float64)[:, :16]
    cols = [torch.ones(g.shape[0], 1, dtype=torch.float64, device=g.device), g]
    for i in range(16):
        cols.append(g[:, i:i + 1] * g[:, :i + 1])
    return torch.cat(cols, dim=1)


def fit_quadratic(dev: Screen, genes, y, alpha: float = 1e-3) -> np.ndarray:
    """FP64 ridge refit of the 153-term surface on the GPU."""
    g = torch.as_tensor(np.asarray(genes, dtype=np.float64)[:, :16])
    return ridge.refit(g, torch.as_tensor(np.asarray(y, dtype=np.float64)), alpha, None, dev).cpu().numpy()


def quad_model(w: np.ndarray) -> Model:
    """The screen model of a 153-term weight vector (f32, as a model file
    written with "%.8e" and read back would give)."""
    return Model.from_compact([float(np.float32(x)) for x in w])


def quad_predict(w: np.ndarray, genes) -> np.ndarray:
    return (features16(genes) @ torch.as_tensor(w, dtype=torch.float64)).numpy()


def train_mlp(dev: Screen, genes, y, epochs: int = 300, batch: int = 4096, lr: float = 1e-3, wd: float = 0.01,
              seed: int = 7):
    """AdamW on the hand-written BF16 forward/backward kernels; returns the
    MLP and its per-epoch training loss (standardized targets)."""
    m = MLP(dev, seed=seed)
    hist = m.fit(torch.tensor(np.asarray(genes, dtype=np.float32)[:, :16]),
                 torch.tensor(np.asarray(y, dtype=np.float32)), epochs=epochs, batch=batch, lr=lr, wd=wd, seed=0)
    return m, hist


def holdout_report(m: MLP, w: np.ndarray, genes, y, k: int = 256) -> Dict[str, float]:
    """Spearman rho and top-k recall against the held-out labels, for the
    MLP (FP64 forward of its weights and BF16 kernel forward) and the
    quadratic surface."""
    g = np.asarray(genes, dtype=np.float32)[:, :16]
    p64 = forward_f64(m.params, g)
    pbf = m.forward(torch.tensor(g)).cpu().numpy()
    pq = quad_predict(w, g)
    return {"n": int(len(y)), "k": k,
            "spearman_mlp": spearman(p64, y), "spearman_mlp_bf16": spearman(pbf, y), "spearman_quad": spearman(pq, y),
            "recall_mlp": topk_recall(p64, y, k), "recall_mlp_bf16": topk_recall(pbf, y, k),
            "recall_quad": topk_recall(pq, y, k)}


def _legacy_rows(surv) -> np.ndarray:
    """Survivor genes 0..15; the label source sees the other genes at their
    legacy values, like every genome of the label set."""
    return np.asarray(surv["g"], dtype=np.float32)[:, :16]


def _next_elites(surv, gen: int):
    from ._abi import Genome
    arr = (Genome * len(surv))()
    for r, s in enumerate(surv):
        e = arr[r]
        e.id, e.gen = gen * 1000 + r, gen + 1
        e.pa, e.pb, e.tribe = int(s["pa"]), int(s["pb"]), int(s["tribe"])
        for k in range(16):
            e.g[k] = float(s["g"][k])
        for k in range(6):
            e.reg[k] = 0.5
    return arr


def funnel_compare(dev: Screen, elites, m: MLP, w: np.ndarray, evaluator, gens: int = 2,
                   lam: int = 2_650_000_000, keep: int = 64, threads: int = 4, scorer: int = SCORER_MLP_BF16,
                   sigma0: float = 0.12, jump: float = 0.01, imm: float = 0.01, pool: int = 1024,
                   progress: Optional[Callable[[Dict[str, float]], None]] = None,
                   train: Optional[tuple] = None, replicate: int = 4,
                   mlp_epochs: int = 300) -> Dict[str, object]:
    """Three screens from the same elites for `gens` chained generations
    (survivors become that screen's next elites, genes 16..28 reset to the
    legacy values); each generation's `keep` survivors are re-scored by the
    codec label source:
      quad      the quadratic surface screen (FP32 on the device);
      mlp       the BF16 MLP screen, its top 4 keep re-ranked in FP64;
      ensemble  the BF16 MLP screen keeps `pool` on the device, re-ranked by
                the mean of the FP64 MLP and quadratic predictions (both
                predict the label).
    Returns per-generation oracle statistics and timings, and per arm the
    mean over generations of the survivors' mean label, the best label seen
    and the last generation's mean.  `progress` is called with each
    generation's record as it completes.

    train = (genes, labels): refit each arm's surrogates after every
    generation on these pairs plus every codec-labelled survivor of that arm
    so far (each repeated `replicate` times), as the reference's loop refits
    its surface on oracle rows (brain/evolve.py:474-478): the quadratic
    surface by the FP64 ridge, the MLP by `mlp_epochs` of AdamW from a fresh
    initialisation.  Without it the surrogates stay fixed."""
    arms = ("quad", "mlp", "ensemble")
    el = {a: elites for a in arms}
    qw = {a: w for a in arms}       # each arm's quadratic weights
    mm = {a: m for a in arms}       # each arm's MLP
    seen = {a: ([], []) for a in arms}  # each arm's labelled survivors
    owned: List[MLP] = []
    out: List[Dict[str, float]] = []
    base = evaluator.labels(_legacy_rows({"g": np.array([list(e.g)[:16] for e in elites])}))
    for gen in range(gens):
        sg = float(np.float32(float("%.4f" % (sigma0 * 0.93 ** gen))))
        cfg = Cfg(lam, keep, gen, threads, sg, float(np.float32(jump)), float(np.float32(imm)))
        rec = {"gen": gen}
        for name in arms:
            t0 = time.perf_counter()
            if name == "quad":
                surv, n = dev.run(quad_model(qw[name]), el[name], cfg)
            elif name == "mlp":
                surv, n = screen_funnel(dev, el[name], cfg, scorer, mm[name], rerank=4)
            else:
                surv, n = screen_funnel(dev, el[name], cfg, scorer, mm[name], rerank=pool // keep,
                                        quad_w=qw[name])
            rec[f"{name}_screen_s"] = time.perf_counter() - t0
            assert n == lam and len(surv) == keep
            t0 = time.perf_counter()
            lab = evaluator.labels(_legacy_rows(surv))
            rec[f"{name}_oracle_s"] = time.perf_counter() - t0
            rec[f"{name}_mean"] = float(lab.mean())
            rec[f"{name}_max"] = float(lab.max())
            rec[f"{name}_top8_mean"] = float(np.sort(lab)[-8:].mean())
            el[name] = _next_elites(surv, gen)
            if train is not None and gen + 1 < gens:
                t0 = time.perf_counter()
                seen[name][0].append(_legacy_rows(surv))
                seen[name][1].append(np.asarray(lab, dtype=np.float64))
                G = np.concatenate([np.asarray(train[0], np.float32)[:, :16]] + seen[name][0] * replicate)
                Y = np.concatenate([np.asarray(train[1], np.float64)] + seen[name][1] * replicate)
                if name in ("quad", "ensemble"):
                    qw[name] = fit_quadratic(dev, G, Y)
                if name in ("mlp", "ensemble"):
                    nm, _ = train_mlp(dev, G, Y, epochs=mlp_epochs)
                    owned.append(nm)
                    mm[name] = nm
                rec[f"{name}_refit_s"] = time.perf_counter() - t0
        out.append(rec)
        if progress is not None:
            progress(rec)
    for nm in owned:
        nm.close()
    summary = {a: {"mean_of_means": float(np.mean([r[f"{a}_mean"] for r in out])),
                   "best": float(max(r[f"{a}_max"] for r in out)),
                   "last_mean": out[-1][f"{a}_mean"]} for a in arms}
    return {"elites_mean": float(base.mean()), "elites_max": float(base.max()), "generations": out,
            "summary": summary}
