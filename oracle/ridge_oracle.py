"""FP64 restatement of the brain's surrogate refit (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's CPU leg use this module,
as the checker of the GPU ridge kernels.

  features        brain/evolve.py:211-220
  load_weights    brain/evolve.py:223-241
  save_weights    brain/evolve.py:244-247
  refit_surrogate brain/evolve.py:250-268  (solve(X'X + a I, X'y + a w_prior))
  loss / grad     the mean ridge loss L(w) = (|Xw - y|^2 + a |w - w_p|^2) / n whose
                  minimizer is refit_surrogate's solution (no reference counterpart)
numpy is the reference's own dependency at this call site (unversioned there).
"""
from __future__ import annotations

import numpy as np


def features(g: np.ndarray) -> np.ndarray:
    """[1, g_0..g_{ng-1}, g_i g_j for i < 16, j <= i]; g is (n, ng) or (ng,)."""
    g = np.asarray(g, dtype=np.float64)
    one = g.ndim == 1
    if one:
        g = g[None, :]
    cols = [np.ones(g.shape[0])] + [g[:, k] for k in range(g.shape[1])]
    for i in range(16):
        for j in range(i + 1):
            cols.append(g[:, i] * g[:, j])
    X = np.stack(cols, axis=1)
    return X[0] if one else X


def refit(X: np.ndarray, y: np.ndarray, alpha: float, prior=None) -> np.ndarray:
    p = X.shape[1]
    w0 = np.zeros(p) if prior is None else np.asarray(prior, dtype=np.float64)
    return np.linalg.solve(X.T @ X + alpha * np.eye(p), X.T @ y + alpha * w0)


def refit_lstsq(X: np.ndarray, y: np.ndarray, alpha: float, prior=None) -> np.ndarray:
    """The same minimizer from the augmented least-squares problem
    [X; sqrt(a) I] w = [y; sqrt(a) w_p] (an independent formulation)."""
    p = X.shape[1]
    w0 = np.zeros(p) if prior is None else np.asarray(prior, dtype=np.float64)
    A = np.vstack([X, np.sqrt(alpha) * np.eye(p)])
    b = np.concatenate([y, np.sqrt(alpha) * w0])
    return np.linalg.lstsq(A, b, rcond=None)[0]


def loss_grad(X, y, w, alpha, prior=None):
    n = X.shape[0]
    w0 = np.zeros_like(w) if prior is None else np.asarray(prior, dtype=np.float64)
    r = X @ w - y
    d = w - w0
    return (r @ r + alpha * d @ d) / n, 2.0 * (X.T @ r + alpha * d) / n


def load_weights(path):
    """brain/evolve.py:223-241: migrate 153/159/161/164/165 to 166."""
    with open(path) as f:
        toks = f.read().split()
    n = int(toks[0])
    w = [float(x) for x in toks[1:1 + n]]
    if n == 153:
        w = w[:17] + [0.0] * 13 + w[17:]
    elif n == 159:
        w = w[:17] + [0.0] * 7 + w[17:]
    elif n == 161:
        w = w[:19] + [0.0] * 5 + w[19:]
    elif n == 164:
        w = w[:22] + [0.0] * 2 + w[22:]
    elif n == 165:
        w = w[:29] + [0.0] + w[29:]
    return w
