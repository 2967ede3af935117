"""bench.py's contract pieces that need no GPU: the one-GPU refusal and the
roofline fractions it derives from the device's measured instruction rates."""
import os
import re
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)

import bench  # noqa: E402


def _rates(add_ns=1.25, simds=1024):
    ns = {"v_add_u32": add_ns, "v_mul_lo_u32": 1.5 * add_ns, "v_mul_hi_u32": 1.5 * add_ns,
          "v_mad_u64_u32": 2.0 * add_ns, "v_lshl_add_u64": 1.6 * add_ns, "v_lshrrev_b64": 1.6 * add_ns,
          "v_mfma_f32_16x16x32_bf16": 11.75, "mfma+4pk_fma_f16": 10.5, "v_pk_fma_f16": 1.8,
          "v_mfma_f32_16x16x128_f8f6f4": 13.5}
    return {"ns_per_instr": ns, "simds": simds}


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stdout + r.stderr


def test_valu_fraction_is_one_at_the_probe_rate():
    rates = _rates()
    add_rate = 1024 * 64 / 1.25e-9
    m = bench.valu_model(rates, add_rate / bench.VALU_LANE_INSTR)
    assert m["frac_of_add_stream"] == pytest.approx(1.0, rel=1e-12)
    # a device whose adds cost twice as long halves the sustainable rate
    m2 = bench.valu_model(_rates(add_ns=2.5), add_rate / bench.VALU_LANE_INSTR)
    assert m2["frac_of_add_stream"] == pytest.approx(2.0, rel=1e-12)
    assert m["cost_vs_add"]["v_mad_u64_u32"] == 2.0
    assert m["cost_vs_add"]["v_mul_lo_u32"] == 1.5


def test_valu_classes_are_a_part_of_the_candidate():
    assert sum(bench.VALU_CLASS_PER_CANDIDATE.values()) < bench.VALU_LANE_INSTR
    m = bench.valu_model(_rates(), 1e10)
    assert m["instr_per_candidate"]["total"] == bench.VALU_LANE_INSTR


@pytest.mark.parametrize("scorer,ins,flop", [(1, "v_mfma_f32_16x16x32_bf16", 16384),
                                             (2, "v_mfma_f32_16x16x128_f8f6f4", 65536)])
def test_mfma_fraction_against_the_layer2_instruction(scorer, ins, flop):
    rates = _rates()
    probe_flops = 1024 * flop / (rates["ns_per_instr"][ins] * 1e-9)
    m = bench.mfma_model(rates, probe_flops, scorer)
    assert m["probe_instruction"] == ins
    assert m["frac_of_probe"] == pytest.approx(1.0, rel=1e-12)
    assert m["ns_4pk"] == pytest.approx(4 * 1.8)


def test_profiled_constants_describe_the_built_source():
    """roofline.traffic and the VALU instruction count are copied from the
    rocprof summary in profiles/; they hold only for the kernel source that
    was profiled, so a change to that source must come with a new profile
    (and the new hash here and in the summary)."""
    assert bench.source_sha(bench.PROFILED_SOURCE) == bench.PROFILED_SHA16
    with open(os.path.join(REPO, "profiles", "screen_walk_offsets_2p65e9.md")) as f:
        summary = f.read()
    assert f"sha256 prefix `{bench.PROFILED_SHA16}`" in summary
    row = re.search(r"\| `k_score` \| SQ_INSTS_VALU \| ([0-9.e+]+) \| ([0-9.]+) \|", summary)
    assert row and round(float(row.group(2))) == bench.VALU_LANE_INSTR
    assert float(row.group(1)) * 64 / 2.65e9 == pytest.approx(float(row.group(2)), abs=0.1)
    assert f"{bench.SCORE_TRAFFIC_PER_LAUNCH / 1e6:.2f} MB per\nlaunch" in summary


def test_next_elites_fills_every_field():
    # the chained generations' elites: ids, generation, lineage, tribe, the
    # 23 genes and 6 regulators of each survivor, in cs_genome's layout
    sys.path.insert(0, os.path.join(REPO, "codec-arena_amd"))
    import numpy as np
    from csfunnel import screen as scr
    rng = np.random.default_rng(4)
    surv = np.zeros(128, dtype=scr.SURVIVOR_DTYPE)
    surv["g"] = rng.random((128, 29)).astype(np.float32)
    surv["pa"] = rng.integers(0, 2 ** 63, 128)
    surv["pb"] = rng.integers(0, 2 ** 63, 128)
    surv["tribe"] = rng.integers(0, 12, 128)
    el = bench.next_elites(surv, 6)
    assert len(el) == 128
    for r in (0, 1, 77, 127):
        e = el[r]
        assert (e.id, e.gen, e.pa, e.pb, e.tribe) == (6 * 128 + r, 7, int(surv["pa"][r]), int(surv["pb"][r]),
                                                      int(surv["tribe"][r]))
        assert list(e.g) == [float(x) for x in surv["g"][r, :23]]
        assert list(e.reg) == [float(x) for x in surv["g"][r, 23:]]
