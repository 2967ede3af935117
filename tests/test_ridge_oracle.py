"""Pins the FP64 ridge oracle (oracle/ridge_oracle.py): feature order against
the screen's model layout, the brain's weight migration against
Model.fromCompact, two independent solvers, and the loss gradient."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import ridge_oracle as ro  # noqa: E402


def test_features_known_values():
    g = np.arange(29) / 100.0
    x = ro.features(g)
    assert x.shape == (166,)
    assert x[0] == 1.0 and np.array_equal(x[1:30], g)
    # quadratic block: (i, j<=i) with i outer: g0g0, g1g0, g1g1, g2g0, ...
    assert x[30] == 0.0 and x[31] == 0.01 * 0.0 and x[32] == 0.01 * 0.01 and x[34] == 0.02 * 0.01
    assert x[30 + 135] == 0.15 * 0.15
    assert ro.features(g[:16]).shape == (153,)


def test_features_match_screen_model_layout(oracle):
    # phi(g).w + 15 w0 is the screen's score of g under Model.fromCompact(w)
    rng = np.random.default_rng(4)
    w = rng.standard_normal(153).astype(np.float32)
    m = oracle.model_from_compact(w)
    for _ in range(50):
        g = rng.random(29).astype(np.float32)
        s64 = ro.features(g[:16].astype(np.float64)) @ w.astype(np.float64) + 15 * float(w[0])
        assert abs(oracle.score(m, g) - s64) < 1e-4 * (1 + abs(s64))


@pytest.mark.parametrize("n", [153, 159, 161, 164, 165])
def test_weight_migration_matches_from_compact(oracle, tmp_path, n):
    # brain/evolve.py:223-241 migrates a model to 166 terms; the migrated model
    # must be the same flat model (src/screen.zig:36-95), except for the 159
    # layout whose regulators the screen reads as genes 16..21 (R4)
    w = (np.arange(n) + 1).astype(np.float32)
    p = tmp_path / "m.txt"
    p.write_text(f"{n}\n" + " ".join(f"{x:.8e}" for x in w) + "\n")
    w166 = np.array(ro.load_weights(str(p)), dtype=np.float32)
    assert w166.shape == (166,)
    a, b = oracle.model_from_compact(w), oracle.model_from_compact(w166)
    assert bytes(a)[4 * 30:] == bytes(b)[4 * 30:]  # the quadratic block
    if n == 159:
        assert list(a.lin[16:22]) == list(w[17:23]) and list(b.lin[23:29]) == list(w[17:23])
    elif n == 165:
        # brain/evolve.py:238-239 inserts the sb16 zero at index 29, after the
        # regulator weights: regulator 0 lands on sb16, the rest shift by one
        assert list(b.lin[:22]) == list(a.lin[:22])
        assert b.lin[22] == w[23] and list(b.lin[23:28]) == list(w[24:29]) and b.lin[28] == 0
    else:
        assert bytes(a) == bytes(b)


def test_two_solvers_agree():
    rng = np.random.default_rng(0)
    X = ro.features(rng.random((800, 16)))
    y = X @ rng.standard_normal(153) + rng.normal(0, 0.5, 800)
    for alpha, prior in ((1e-3, None), (50.0, rng.standard_normal(153))):
        a = ro.refit(X, y, alpha, prior)
        b = ro.refit_lstsq(X, y, alpha, prior)
        assert np.max(np.abs(a - b)) < 1e-6 * (1 + np.max(np.abs(a)))


def test_loss_gradient_finite_differences_and_minimum():
    rng = np.random.default_rng(1)
    X = ro.features(rng.random((500, 29)))
    y = rng.standard_normal(500)
    prior = rng.standard_normal(166)
    w = rng.standard_normal(166)
    L, g = ro.loss_grad(X, y, w, 0.3, prior)
    h = 1e-4
    for k in (0, 5, 29, 100, 165):
        e = np.zeros(166)
        e[k] = h
        fd = (ro.loss_grad(X, y, w + e, 0.3, prior)[0] - ro.loss_grad(X, y, w - e, 0.3, prior)[0]) / (2 * h)
        assert abs(fd - g[k]) < 1e-7 * (1 + abs(g[k]))
    ws = ro.refit(X, y, 0.3, prior)
    _, g0 = ro.loss_grad(X, y, ws, 0.3, prior)
    assert np.max(np.abs(g0)) < 1e-8


def test_save_weights_round_trips_f32(tmp_path):
    from csfunnel import ridge
    w = np.random.default_rng(2).standard_normal(166).astype(np.float32)
    p = str(tmp_path / "w.txt")
    ridge.save_weights(p, [float(x) for x in w])
    toks = open(p).read().split()
    assert toks[0] == "166" and len(toks) == 167
    assert np.array_equal(np.array([np.float32(t) for t in toks[1:]]), w)
    assert ridge.load_weights(p) == ro.load_weights(p)
