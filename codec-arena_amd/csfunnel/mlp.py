"""Learned surrogate: a 16 -> 256 -> 256 -> 1 GELU MLP over the 16 legacy
effector genes, trained on the GPU by libcsfunnel's BF16 MFMA forward and
backward kernels and used as a screen scorer in BF16 or FP8.

No counterpart in the reference (codec-arena ranks candidates with the
quadratic surface of src/screen.zig:97-112 only); the training targets are
(genome, score) pairs from an oracle.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Tuple

import numpy as np
import torch

from ._abi import Genome, Survivor, check, lib
from .screen import SURVIVOR_DTYPE, Cfg, Model, Screen

N_PARAMS = 70401
IN, H = 16, 256
SCORER_QUAD, SCORER_MLP_BF16, SCORER_MLP_FP8, SCORER_MLP_F64 = 0, 1, 2, 3


class ScreenOpts(C.Structure):
    _fields_ = [("scorer", C.c_int), ("_pad", C.c_int), ("mlp", C.c_void_p), ("dump_genes", C.c_void_p),
                ("dump_scores", C.c_void_p), ("scores64", C.c_void_p)]


def unpack(p: np.ndarray):
    """Flat parameters -> (W1 [16,256], b1, W2 [256,256], b2, w3, b3)."""
    o = 0
    W1 = p[o:o + IN * H].reshape(IN, H); o += IN * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def _f32(t) -> torch.Tensor:
    return torch.as_tensor(t, dtype=torch.float32, device="cuda").contiguous()


class MLP:
    def __init__(self, dev: Screen, seed: int = 0):
        self.dev = dev
        h = C.c_void_p()
        check(lib().cs_mlp_create(dev._h, seed, C.byref(h)))
        self._h = h
        dev._children.add(self)

    def close(self):
        if self._h:
            lib().cs_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self) -> np.ndarray:
        p = np.zeros(N_PARAMS, dtype=np.float32)
        check(lib().cs_mlp_get_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float))))
        return p

    @params.setter
    def params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == N_PARAMS
        check(lib().cs_mlp_set_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float))))

    def forward(self, genes) -> torch.Tensor:
        g = _f32(genes)
        out = torch.empty(g.shape[0], dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        check(lib().cs_mlp_forward(self._h, C.c_void_p(g.data_ptr()), g.shape[1], g.shape[0],
                                   C.c_void_p(out.data_ptr()), None))
        return out

    def h1_max(self, genes) -> float:
        g = _f32(genes)
        out = torch.empty(g.shape[0], dtype=torch.float32, device="cuda")
        hm = C.c_float()
        torch.cuda.synchronize()
        check(lib().cs_mlp_forward(self._h, C.c_void_p(g.data_ptr()), g.shape[1], g.shape[0],
                                   C.c_void_p(out.data_ptr()), C.byref(hm)))
        return hm.value

    def grad(self, genes, y) -> Tuple[float, np.ndarray]:
        g, t = _f32(genes), _f32(y)
        loss = C.c_float()
        gr = np.zeros(N_PARAMS, dtype=np.float32)
        torch.cuda.synchronize()
        check(lib().cs_mlp_grad(self._h, C.c_void_p(g.data_ptr()), g.shape[1], C.c_void_p(t.data_ptr()),
                                g.shape[0], C.byref(loss), gr.ctypes.data_as(C.POINTER(C.c_float))))
        return loss.value, gr

    def train_step(self, genes, y, lr: float = 1e-3, wd: float = 0.01) -> float:
        g, t = _f32(genes), _f32(y)
        loss = C.c_float()
        torch.cuda.synchronize()
        check(lib().cs_mlp_train_step(self._h, C.c_void_p(g.data_ptr()), g.shape[1], C.c_void_p(t.data_ptr()),
                                      g.shape[0], C.c_float(lr), C.c_float(wd), C.byref(loss)))
        return loss.value

    def calibrate_fp8(self, genes, headroom: float = 224.0) -> float:
        """The E4M3 scale of H1, layer 2's activations: the largest power of
        two that keeps max |H1| over `genes` within `headroom` (half of
        E4M3's 448, leaving room for candidates beyond the sample).  W2 gets
        per-block E8M0 scales from the library."""
        hmax = max(self.h1_max(genes), 1e-12)
        hs = float(2.0 ** min(15, max(-15, int(np.floor(np.log2(headroom / hmax))))))
        check(lib().cs_mlp_set_fp8_scale(self._h, C.c_float(hs)))
        return hs

    def fit(self, genes, y, epochs: int, batch: int = 4096, lr: float = 1e-3, wd: float = 0.01, seed: int = 0,
            target_norm: bool = True):
        """AdamW on shuffled mini-batches; targets are standardized when
        target_norm (the affine map is folded back into w3, b3 at the end)."""
        g = _f32(genes)
        t = _f32(y)
        mu, sd = (float(t.mean()), float(t.std())) if target_norm else (0.0, 1.0)
        tn = (t - mu) / sd
        n = g.shape[0]
        gen = torch.Generator(device="cpu").manual_seed(seed)
        hist = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=gen).cuda()
            tot = 0.0
            for s in range(0, n, batch):
                idx = perm[s:s + batch]
                tot += self.train_step(g[idx], tn[idx], lr, wd) * idx.numel()
            hist.append(tot / n)
        if target_norm:
            p = self.params
            W1, b1, W2, b2, w3, b3 = unpack(p)
            w3 *= sd
            p[-1] = b3 * sd + mu
            self.params = p
        return hist


def screen_ex(dev: Screen, model: Optional[Model], elites, cfg: Cfg, scorer: int, mlp: Optional[MLP] = None,
              dump: bool = False, keys: Optional[np.ndarray] = None):
    """cs_screen_ex: survivors, screens_count and (dump) every candidate's
    genes [lambda, 29] and score [lambda] as cuda tensors in stream order.
    keys (float64 [keep]) receives each survivor's ranking key: the FP64
    score for SCORER_MLP_F64."""
    ne = len(elites)
    earr = elites if isinstance(elites, C.Array) else (Genome * max(1, ne))(*elites)
    cap = max(1, int(cfg.keep))
    out = np.zeros(cap, dtype=SURVIVOR_DTYPE)
    n_out, screens = C.c_uint64(), C.c_uint64()
    dg = ds = None
    if dump:
        dg = torch.empty((cfg.lambda_, 29), dtype=torch.float32, device="cuda")
        ds = torch.empty(cfg.lambda_, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
    if keys is not None:
        assert keys.dtype == np.float64 and keys.flags.c_contiguous and keys.size >= cap
    opts = ScreenOpts(scorer, 0, mlp._h if mlp is not None else None, dg.data_ptr() if dump else None,
                      ds.data_ptr() if dump else None, keys.ctypes.data if keys is not None else None)
    c = cfg.to_c()
    check(lib().cs_screen_ex(dev._h, C.byref(model.raw) if model is not None else None, earr, ne, C.byref(c),
                             C.byref(opts), out.ctypes.data_as(C.POINTER(Survivor)), cap, C.byref(n_out),
                             C.byref(screens)))
    return out[: n_out.value].copy(), screens.value, dg, ds


def screen_f64(dev: Screen, elites, cfg: Cfg, mlp: MLP):
    """The exact top `keep` of the MLP evaluated in FP64 on the device over
    every candidate of the generation (no dump): survivors, their FP64 scores
    and screens_count."""
    keys = np.zeros(max(1, int(cfg.keep)), dtype=np.float64)
    surv, screens, _, _ = screen_ex(dev, None, elites, cfg, SCORER_MLP_F64, mlp, keys=keys)
    return surv, keys[: len(surv)].copy(), screens


def forward_f64(params: np.ndarray, genes: np.ndarray) -> np.ndarray:
    """The MLP in FP64 on the host (FP32 master weights), the re-ranker and
    the reference the screens' recall is measured against."""
    W1, b1, W2, b2, w3, b3 = [np.asarray(t, dtype=np.float64) for t in unpack(params)]
    x = np.asarray(genes, dtype=np.float64)[:, :IN]

    def gelu(z):
        # the cube as products, in the device FP64 scorer's order (numpy's
        # z ** 3 goes through pow and costs ~80x more)
        return 0.5 * z * (1.0 + np.tanh(0.7978845608028654 * (z + 0.044715 * z * z * z)))

    return gelu(gelu(x @ W1 + b1) @ W2 + b2) @ w3 + float(b3)


def screen_funnel(dev: Screen, elites, cfg: Cfg, scorer: int, mlp: MLP, rerank: int = 4,
                  quad_w: Optional[np.ndarray] = None):
    """The full funnel: the device screen keeps rerank * keep candidates by
    the BF16 / FP8 score, the host re-scores those in FP64 and returns the top
    `keep` (scores replaced by the FP64 values, FP64 order).  With quad_w (a
    153-term surface over genes 0..15 predicting the same label) the re-rank
    score is the mean of the two FP64 predictions."""
    wide = Cfg(cfg.lambda_, cfg.keep * rerank, cfg.seed, cfg.threads, cfg.sigma, cfg.jump_p, cfg.imm_p)
    surv, screens, _, _ = screen_ex(dev, None, elites, wide, scorer, mlp)
    s64 = forward_f64(mlp.params, surv["g"])
    if quad_w is not None:
        from .surrogate import quad_predict
        s64 = 0.5 * (s64 + quad_predict(quad_w, np.asarray(surv["g"], dtype=np.float32)[:, :16]))
    order = np.argsort(-s64, kind="stable")[: cfg.keep]
    out = surv[order].copy()
    out["score"] = s64[order].astype(np.float32)
    return out, screens


def spearman(a, b) -> float:
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    ra = np.argsort(np.argsort(a, kind="stable"), kind="stable").astype(np.float64)
    rb = np.argsort(np.argsort(b, kind="stable"), kind="stable").astype(np.float64)
    ra -= ra.mean()
    rb -= rb.mean()
    return float(ra @ rb / np.sqrt((ra @ ra) * (rb @ rb)))


def topk_recall(pred, truth, k: int) -> float:
    """|top-k(pred) & top-k(truth)| / k."""
    pi = set(np.argsort(-np.asarray(pred), kind="stable")[:k].tolist())
    ti = set(np.argsort(-np.asarray(truth), kind="stable")[:k].tolist())
    return len(pi & ti) / k
