"""cs_screen_noalloc (device buffers from the library's static scratch array)
against cs_screen (hipMalloc'd context buffers): the same survivors, bit for
bit, on consecutive calls that share the one scratch array, the second and
third with smaller inputs than the call before; the transfer-map path; a
full-size generation; and a clean CS_E_CAP for a call beyond the scratch
size."""
import os

import numpy as np
import pytest

from csfunnel import screen as scr
from csfunnel._abi import CsError

pytestmark = pytest.mark.gpu


def _same(a, b):
    assert len(a) == len(b)
    for f in ("score", "g", "pa", "pb", "tribe", "stream", "index", "start"):
        x, y = a[f], b[f]
        if x.dtype.kind == "f":
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), f
        else:
            assert np.array_equal(x, y), f


def _cfg(lam, keep, seed, sigma=0.12, jump=0.01, imm=0.01, threads=4):
    return scr.Cfg(lam, keep, seed, threads, float(np.float32(sigma)), float(np.float32(jump)),
                   float(np.float32(imm)))


@pytest.fixture(scope="module")
def inputs(golden):
    model = scr.load_model(os.path.join(golden, "c0_model.txt"))
    el128 = scr.load_elites(os.path.join(golden, "c1_elites.txt"))
    el64 = scr.load_elites(os.path.join(golden, "c0_elites.txt"))
    return model, el128, el64


def test_noalloc_matches_cs_screen_on_consecutive_calls(inputs):
    model, el128, el64 = inputs
    cases = [(el128, _cfg(1 << 22, 256, 11)),          # first call
             (el64, _cfg(300_000, 100, 12, threads=3)),  # smaller: fewer elites, candidates, survivors
             (el64, _cfg(5000, 7, 13, jump=0.02))]      # smaller again
    with scr.Screen(0) as dev:
        for el, cfg in cases:
            print(f"case noalloc lam={cfg.lambda_} keep={cfg.keep} ne={len(el)} T={cfg.threads}", flush=True)
            ref, rs = dev.run(model, el, cfg)
            got, gs = dev.run(model, el, cfg, noalloc=True)
            assert gs == rs == cfg.lambda_
            _same(got, ref)


def test_noalloc_transfer_maps_and_small_record_buffers(inputs):
    model, el128, _ = inputs
    with scr.Screen(0) as dev:
        dev.set_offset_mode(1)
        cfg = _cfg(1 << 24, 128, 21)
        print(f"case noalloc maps lam={cfg.lambda_}", flush=True)
        ref, _ = dev.run(model, el128, cfg)
        got, _ = dev.run(model, el128, cfg, noalloc=True)
        assert dev.stats()["offset_path"] == 0
        _same(got, ref)
        # record buffers of 2048: overflowing launches are split and redone
        dev.set_offset_mode(0)
        dev.set_record_cap(2048)
        cfg = _cfg(1 << 20, 256, 22)
        print("case noalloc record cap 2048", flush=True)
        got, _ = dev.run(model, el128, cfg, noalloc=True)
        assert dev.stats()["retries"] > 0
        ref, _ = dev.run(model, el128, cfg)
        _same(got, ref)


def test_noalloc_production_generation(inputs):
    model, el128, _ = inputs
    cfg = _cfg(2_650_000_000, 128, 0)
    with scr.Screen(0) as dev:
        print("case noalloc production generation lam=2.65e9", flush=True)
        got, gs = dev.run(model, el128, cfg, noalloc=True)
        assert gs == cfg.lambda_ and dev.stats()["candidates"] == cfg.lambda_
        ref, _ = dev.run(model, el128, cfg)
        _same(got, ref)


def test_noalloc_rejects_calls_beyond_the_scratch(inputs):
    model, el128, _ = inputs
    with scr.Screen(0) as dev:
        dev.set_offset_mode(1)  # transfer maps: 512 bytes per 65536-draw block
        print("case noalloc beyond scratch", flush=True)
        with pytest.raises(CsError) as e:
            dev.run(model, el128, _cfg(1 << 28, 128, 3), noalloc=True)
        assert e.value.name == "Capacity" and "scratch" in str(e.value)
        # the context is still usable, and so is the scratch array
        dev.set_offset_mode(0)
        cfg = _cfg(100_000, 64, 4)
        got, _ = dev.run(model, el128, cfg, noalloc=True)
        ref, _ = dev.run(model, el128, cfg)
        _same(got, ref)
