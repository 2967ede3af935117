"""Stand-in label source for the learned-surrogate rows (TEST INFRASTRUCTURE).

The north star labels genomes with codec-arena's encode-decode oracle
(PSNR - 14 bpp over the synthetic corpus, src/codec.zig:2537-2670,
brain/config.json:2); that codec is not restated in this repository yet.
Until it is, training and evaluation use this rate-distortion model of the
same 16 legacy effector genes.  It maps genes to codec parameters with the
reference's phenotype functions (src/genome.zig:205-252: key interval,
DC/deadzone multiplier, search range, base quantizer, band scales, deadzone,
skip threshold, P-frame q multiplier, intra band-0 boost, dither, half-pel
switch, intra refresh, adaptive q, chroma scale) and scores them with
textbook high-rate quantization: per-band distortion q^2/12 (capped by the
source variance, raised by the deadzone) and rate max(0, log2(sigma/q) + 1/2)
bits per coefficient, intra vs. motion-compensated residual sources, skip
and key-frame fractions, then PSNR - 14 bpp plus N(0, noise) measurement
noise.  It exercises and measures the MLP machinery; results on it are not
claims about the codec.
"""
from __future__ import annotations

import numpy as np

BANDS_N = np.array([4.0, 16.0, 44.0])      # coefficients per 8x8 block in low / mid / high bands
SIG_I = np.array([48.0, 18.0, 7.0])        # intra coefficient std per band
LAMBDA = 14.0                              # brain/config.json:2


def _band_dr(q, sig, n, dz):
    """distortion (sum over the band's coefficients) and rate (bits)."""
    d0 = q * q / 12.0 * (1.0 + 0.08 * dz)
    d = n * np.minimum(sig * sig, d0)
    r = n * np.maximum(0.0, np.log2(sig / q) + 0.5)
    return d, r


class LabelProxy:
    def __init__(self, noise: float = 0.5):
        self.noise = noise

    def clean(self, genes: np.ndarray) -> np.ndarray:
        g = np.clip(np.asarray(genes, dtype=np.float64)[:, :16], 0.0, 1.0)
        keyint = 1.0 + np.floor(g[:, 0] * 47.0)
        dzdc = 0.5 + 1.5 * g[:, 1]
        search = np.floor(g[:, 2] * 12.0)
        baseq = 2.0 + 46.0 * g[:, 3]
        bs = [0.5 + 1.5 * g[:, 4], 0.5 + 2.5 * g[:, 5], 0.5 + 4.0 * g[:, 6]]
        dz = 14.0 * g[:, 7]
        skip = 8.0 * g[:, 8]
        pq = 0.3 + 1.7 * g[:, 9]
        ib0 = 0.5 + 1.5 * g[:, 10]
        dith = g[:, 11] - 0.5
        halfpel = g[:, 12] > 0.6
        refresh = 1.0 + np.floor(8.0 * g[:, 13])
        adapt = 10.0 * g[:, 14]
        chroma = 0.5 + 1.5 * g[:, 15]
        # motion-compensated residual: prediction gain from search range and half-pel
        pgain = 0.45 / (1.0 + 0.12 * search) * np.where(halfpel, 0.82, 1.0)
        dI = rI = dP = rP = 0.0
        for b in range(3):
            qI = baseq * bs[b] * (dzdc * ib0 if b == 0 else 1.0) * 0.5
            qP = baseq * bs[b] * pq * (dzdc if b == 0 else 1.0) * 0.5
            d, r = _band_dr(qI, SIG_I[b], BANDS_N[b], dz)
            dI, rI = dI + d, rI + r
            d, r = _band_dr(qP, SIG_I[b] * pgain, BANDS_N[b], dz)
            dP, rP = dP + d, rP + r
        # skipped P blocks: no residual bits, the residual becomes distortion
        fskip = 1.0 / (1.0 + np.exp(-(skip - 30.0 * pgain) / 1.2))
        resid = (SIG_I * SIG_I * BANDS_N).sum() * pgain * pgain
        dP = (1.0 - fskip) * dP + fskip * resid
        rP = (1.0 - fskip) * (rP + 6.0) + fskip * 1.5  # + motion-vector bits
        # adaptive quantization trades rate for distortion, dither adds noise
        aq = 1.0 - 0.035 * adapt + 0.0025 * adapt * adapt
        fI = np.minimum(1.0, 1.0 / keyint + 0.12 / refresh)
        mse = (fI * dI * aq + (1.0 - fI) * dP * aq) / 64.0 + 2.0 * dith * dith
        bits = fI * rI * (1.0 + 0.02 * adapt) + (1.0 - fI) * rP
        qc = baseq * chroma * 0.5
        mse += 0.25 * np.minimum(100.0, qc * qc / 12.0) * 0.3
        bits += 0.5 * 64.0 * np.maximum(0.0, np.log2(12.0 / qc) + 0.5) * 0.1
        psnr = 10.0 * np.log10(255.0 ** 2 / np.maximum(mse, 1e-3))
        bpp = bits / 64.0 + 0.01
        return psnr - LAMBDA * bpp

    def __call__(self, genes: np.ndarray, seed: int) -> np.ndarray:
        y = self.clean(genes)
        return y + np.random.default_rng(seed).normal(0.0, self.noise, y.shape[0])
