"""The VALU instruction-rate probe behind the quadratic screen's roofline
(no reference counterpart): every class times to a positive, finite issue
cost on the device, and an add issues no slower than a 32-bit multiply."""
import math

import pytest


@pytest.mark.gpu
def test_valu_rate_probe_measures_every_class():
    from csfunnel import screen as scr
    r = scr.valu_rates(0, iters=1 << 12)
    ns = r["ns_per_instr"]
    print({k: round(v, 4) for k, v in ns.items()}, {k: round(v, 3) for k, v in r["memtime_per_instr"].items()},
          r["memtime_ghz"], r["simds"], flush=True)
    assert r["simds"] >= 4 and set(ns) == set(scr.VALU_CLASSES)
    assert all(math.isfinite(v) and v > 0 for v in ns.values())
    assert 0.01 < r["memtime_ghz"] < 10.0
    assert ns["v_add_u32"] <= ns["v_mul_lo_u32"] * 1.05


def test_valu_rate_probe_rejects_bad_arguments():
    import ctypes as C
    from csfunnel import _abi
    L = _abi.lib()
    ns = C.c_double(0.0)
    assert L.cs_valu_rate_probe(0, 99, 16, C.byref(ns), None, None) < 0
    assert L.cs_valu_rate_probe(0, 0, 0, C.byref(ns), None, None) < 0
    assert L.cs_valu_rate_probe(0, 0, 16, None, None, None) < 0
