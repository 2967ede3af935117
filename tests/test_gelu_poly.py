"""The packed-f16 GELU of the MLP screen (codec-arena_amd/csrc/mlp_common.h):
the coefficients compiled into the kernel are the ones gelu_poly_fit.py
derives, and an f16 emulation of the kernel's evaluation sequence stays
within the stated bounds of the tanh-form GELU (the FP64 reference every
recall figure is measured against)."""
import os
import re
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, "codec-arena_amd", "csrc")
sys.path.insert(0, CSRC)
import gelu_poly_fit as fit  # noqa: E402


def header_coefficients():
    txt = open(os.path.join(CSRC, "mlp_common.h")).read()
    body = txt[txt.index("V gelu_sig_scaled(V v)"):]
    body = body[:body.index("return")]
    vals = [float(v) for v in re.findall(r"\(V\)\(_Float16\)(-?[0-9.]+)f", body)]
    # one, the leading coefficient, the Horner constants, 1/2, 0
    horner = [vals[1]] + vals[2:2 + fit.DEG]
    return np.array(horner[::-1])  # c0 .. cDEG


def test_kernel_coefficients_are_the_fit():
    coef = fit.fit()
    got = header_coefficients()
    assert len(got) == fit.DEG + 1
    assert np.array_equal(np.float16(got), np.float16(coef))


def test_f16_sequence_error_bounds():
    coef = header_coefficients()
    x = np.linspace(-16.0, 16.0, 320001)
    s16, x16 = fit.s_f16(x, coef)
    s_err = np.max(np.abs(s16 - fit.s_exact(x)))
    g16 = fit.h(x16 * s16)
    g_err = np.max(np.abs(g16 - x * fit.s_exact(x)))
    assert s_err < 1.2e-3, s_err
    assert g_err < 8e-3, g_err  # f16 resolution of x near |x| = 16
    # the sigmoid factor saturates correctly far out, with w unclamped: over
    # the whole f16 range of x (and beyond it, where the conversion toward
    # zero stops at 65504) the error bound holds and nothing is NaN
    s_far, _ = fit.s_f16(np.array([-60.0, -8.0, 8.0, 60.0]), coef)
    assert np.allclose(s_far, [0.0, 0.0, 1.0, 1.0], atol=1e-3)
    xf = np.concatenate([np.linspace(-70000.0, 70000.0, 280001), [-1e9, -65504.0, 65504.0, 1e9]])
    sf, _ = fit.s_f16(xf, coef)
    assert not np.isnan(sf).any()
    assert np.max(np.abs(sf - fit.s_exact(xf))) < 1.2e-3
