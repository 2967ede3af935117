"""The library and torch must share one HIP runtime in either import order."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import os, sys
sys.path.insert(0, os.path.join({repo!r}, "codec-arena_amd"))
order = sys.argv[1]
if order == "torch-first":
    import torch
    x = torch.ones(4, device="cuda")
from csfunnel import screen as scr
g = os.path.join({repo!r}, "tests", "golden")
with scr.Screen(0) as d:
    surv, n = d.run(scr.load_model(os.path.join(g, "c0_model.txt")), scr.load_elites(os.path.join(g, "c0_elites.txt")),
                    scr.Cfg(100000, 16, 1))
    assert n == 100000 and len(surv) == 16
    import torch
    t = torch.arange(10, device="cuda", dtype=torch.float64).sum().item()
    assert t == 45.0
maps = open("/proc/self/maps").read()
libs = sorted(set(l.split()[-1] for l in maps.splitlines() if "libamdhip64" in l or "libhsa-runtime64" in l))
assert len([x for x in libs if "amdhip" in x]) == 1 and len([x for x in libs if "hsa-runtime" in x]) == 1, libs
print("ok", order, libs)
"""


@pytest.mark.parametrize("order", ["lib-first", "torch-first"])
def test_one_runtime_either_order(order):
    print(f"case import order {order}", flush=True)
    r = subprocess.run([sys.executable, "-c", SCRIPT.format(repo=REPO), order], capture_output=True, text=True,
                       timeout=240)
    print(r.stdout, r.stderr[-2000:], flush=True)
    assert r.returncode == 0 and "ok" in r.stdout
