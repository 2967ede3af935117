"""csfunnel -- MI355X (gfx950) screening funnel for codec-arena.

Drop-in for the ``arena screen`` call of codec-arena's brain/evolve.py
(brain/evolve.py:431-454): same genome encoding (src/genome.zig), same survivor
list and screens_count (src/screen.zig), computed by hand-written HIP kernels
behind the C ABI in include/csfunnel.h.
"""
from ._abi import CsError, NG, NR, NTOT  # noqa: F401
from .screen import Cfg, Model, Screen, load_elites, load_model, parse_line, screen_cmd  # noqa: F401
