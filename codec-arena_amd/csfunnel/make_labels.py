"""Label set of the learned surrogate: (genome, PSNR - 14 bpp) pairs.

25,000 genomes (20,000 train + 5,000 held-out): the 16 legacy effector genes
uniform in [0, 1) on the 4-decimal lattice that pop files carry
(brain/evolve.py:104-106), genes 16-22 at 0 and regulators at 0.5 (a v1 pop
line, src/genome.zig:142-185), each evaluated by the codec label source on the
ten-clip synthetic corpus.  Work is saved in chunks, so an interrupted run
resumes where it stopped.

    python -m csfunnel.make_labels [--out PATH] [--threads T] [--n 25000]
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

SEED = 3
N_TRAIN, N_HELD = 20000, 5000
DEFAULT_OUT = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "data",
                           "codec_labels_c3.npz")


def genomes(n: int = N_TRAIN + N_HELD, seed: int = SEED) -> np.ndarray:
    """[n, 16] genes on the 4-decimal lattice, as integers 0..9999."""
    return np.round(np.random.default_rng(seed).random((n, 16)) * 10000.0).astype(np.uint16)


def make(out: str, n: int, threads: int, chunk: int = 500) -> None:
    from .codec import Evaluator
    gi = genomes(n)
    parts = out + ".parts"
    os.makedirs(parts, exist_ok=True)
    ev = Evaluator(threads=threads)
    cols = ("psnr", "bpp", "dec_mops", "psnr_u", "psnr_v", "bytes", "label")
    for s in range(0, n, chunk):
        path = os.path.join(parts, f"{s:06d}.npz")
        if os.path.exists(path):
            continue
        t0 = time.time()
        ms = ev.evaluate(gi[s:s + chunk].astype(np.float32) / np.float32(10000.0), ids=range(s, s + chunk))
        np.savez(path, **{c: np.array([m[c] for m in ms]) for c in cols})
        print(f"labels {s + len(ms)}/{n} ({time.time() - t0:.1f} s)", flush=True)
    res = {c: [] for c in cols}
    for s in range(0, n, chunk):
        with np.load(os.path.join(parts, f"{s:06d}.npz")) as z:
            for c in cols:
                res[c].append(z[c])
    np.savez_compressed(out, genes_e4=gi, seed=SEED, n_train=N_TRAIN, clips=np.array(ev.names),
                        **{c: np.concatenate(v) for c, v in res.items()})
    print(f"wrote {out}", flush=True)


def load(path: str = DEFAULT_OUT):
    """(genes [n,16] float32, labels [n] float64, n_train, full record)."""
    z = np.load(path, allow_pickle=False)
    genes = z["genes_e4"].astype(np.float32) / np.float32(10000.0)
    return genes, z["label"].astype(np.float64), int(z["n_train"]), z


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--n", type=int, default=N_TRAIN + N_HELD)
    a = ap.parse_args(argv)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    make(a.out, a.n, a.threads)
    return 0


if __name__ == "__main__":
    sys.exit(main())
