"""Writes the seeded input fixtures of the screen configurations.

    python tests/golden/make_fixtures.py

c0_elites.txt  64 v1 pop lines (21 tokens): id=i gen=0 pa=0 pb=0 tribe=i%12,
               genes g0..g15 = splitmix64(seed 0) units rounded to 4 decimals,
               elite-major (the 16 effectors the 153-term surface scores).
c1_elites.txt  the same generator for 128 elites.
c0_model.txt   153 coefficients ~ N(0,1)/sqrt(153) from numpy default_rng(0),
               written like brain/evolve.py:244-247 save_weights ("%.8e").

surrogate_161.txt and seed_pop.txt are data files of the reference snapshot
(state/surrogate.txt, pop/seed.txt), copied unchanged.
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GAMMA = 0x9E3779B97F4A7C15
M64 = (1 << 64) - 1


def splitmix_units(seed: int, n: int):
    s = seed
    out = []
    for _ in range(n):
        s = (s + GAMMA) & M64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        z ^= z >> 31
        out.append((z >> 40) / 16777216.0)
    return out


def elites(n: int) -> str:
    u = splitmix_units(0, 16 * n)
    lines = []
    for i in range(n):
        genes = " ".join(f"{x:.4f}" for x in u[16 * i:16 * i + 16])
        lines.append(f"{i} 0 0 0 {i % 12} {genes}")
    return "\n".join(lines) + "\n"


def main():
    with open(os.path.join(HERE, "c0_elites.txt"), "w") as f:
        f.write(elites(64))
    with open(os.path.join(HERE, "c1_elites.txt"), "w") as f:
        f.write(elites(128))
    w = np.random.default_rng(0).standard_normal(153) / np.sqrt(153.0)
    with open(os.path.join(HERE, "c0_model.txt"), "w") as f:
        f.write("153\n" + " ".join(f"{x:.8e}" for x in w) + "\n")

if __name__ == "__main__":
    main()
