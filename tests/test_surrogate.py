"""CPU checks of the learned-surrogate host code: the 153-term feature layout
against the oracle's features (brain/evolve.py:211-220) and the screen model
built from refit weights (src/screen.zig:36-112), and the committed codec
label set against the label source it was made with."""
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "oracle"))
import ridge_oracle  # noqa: E402


def test_features16_match_oracle():
    from csfunnel import surrogate
    g = np.random.default_rng(4).random((50, 16))
    assert np.array_equal(surrogate.features16(g).numpy(), ridge_oracle.features(g))
    assert surrogate.features16(g).shape == (50, surrogate.N_QUAD)


def test_quad_model_scores_like_the_refit_weights(oracle):
    """Model.fromCompact of the weights scores a genome as phi(g).w with the
    intercept counted 16 times (R7), within f32 rounding."""
    from csfunnel import surrogate
    w = np.random.default_rng(5).standard_normal(153) / np.sqrt(153.0)
    m = surrogate.quad_model(w)
    g = np.random.default_rng(6).random((200, 16))
    genes = np.zeros((200, 29), dtype=np.float32)
    genes[:, :16] = g
    got = np.array([oracle.score(oracle.model_from_compact([float(np.float32(x)) for x in w]), gg) for gg in genes])
    ref = surrogate.quad_predict(w.astype(np.float32).astype(np.float64), genes[:, :16]) + 15.0 * float(np.float32(w[0]))
    assert np.max(np.abs(got - ref)) < 1e-5 * (1 + np.abs(ref).max())
    assert abs(m.w0 - float(np.float32(w[0]))) == 0.0


LABELS = os.path.join(REPO, "codec-arena_amd", "data", "codec_labels_c3.npz")


@pytest.mark.skipif(not os.path.exists(LABELS), reason="label set not generated")
def test_label_set_reproduces_from_the_label_source():
    """The committed labels are the codec label source's: the genes are the
    seeded draw, and a spread of entries recompute bit-identically."""
    from csfunnel import codec, make_labels
    genes, y, ntr, z = make_labels.load(LABELS)
    assert genes.shape == (25000, 16) and ntr == 20000
    assert np.array_equal(z["genes_e4"], make_labels.genomes())
    assert tuple(z["clips"]) == codec.Evaluator.__init__.__defaults__[0]
    idx = [0, 1, 7777, 19999, 20000, 24999]
    ms = codec.Evaluator().evaluate(genes[idx], ids=idx)
    for i, m in zip(idx, ms):
        assert np.float32(m["psnr"]) == np.float32(z["psnr"][i])
        assert m["bpp"] == z["bpp"][i]
        assert m["label"] == y[i]
    assert np.all(np.isfinite(y)) and y.std() > 1.0
