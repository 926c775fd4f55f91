"""Synthetic filterbank inputs of the benchmark workloads (numpy, host side).

Gaussian noise quantised to unsigned nbits-bit samples, plus dispersed boxcar
pulses placed at the TDD integer delays rn_even(DM * delay[c]) of the plan's
delay table, so the matching TDD trial adds them up exactly.
"""
from __future__ import annotations

import numpy as np


def tdd_delays(dm: float, delays: np.ndarray) -> np.ndarray:
    """rn_even(float32(dm) * delay[c]) as unsigned (negative -> 0)."""
    p = (np.float32(dm) * np.asarray(delays, np.float32)).astype(np.float32)
    r = np.rint(p)
    return np.where(p > 0, r, 0).astype(np.int64)


def filterbank(nsamps: int, nchans: int, nbits: int, mean: float, sigma: float, seed: int,
               pulses=(), delays: np.ndarray | None = None) -> np.ndarray:
    """Unpacked samples uint32 [nsamps, nchans]: clip(round(N(mean, sigma)), 0, 2^nbits-1)
    plus pulses [(dm, t0, width, amplitude), ...] (clipped to the range)."""
    rng = np.random.default_rng(seed)
    vmax = (1 << nbits) - 1 if nbits < 32 else 0xFFFFFFFF
    x = rng.normal(mean, sigma, size=(nsamps, nchans)).astype(np.float32)
    s = np.clip(np.rint(x), 0, vmax).astype(np.int64)
    for dm, t0, width, amp in pulses:
        d = tdd_delays(dm, delays)
        for w in range(int(width)):
            t = t0 + d + w
            ok = t < nsamps
            c = np.nonzero(ok)[0]
            s[t[ok], c] = np.minimum(s[t[ok], c] + int(amp), vmax)
    return s.astype(np.uint32)


# Workload geometries (the BASELINE configurations).
CFG1 = dict(nchans=256, f0=1581.8, df=-0.390625, dt=64e-6, nsamps=65536, nbits=8,
            mean=128.0, sigma=20.0, dms=("linear", 0.0, 1.0, 256))
CFG2 = dict(nchans=1024, f0=1530.0, df=-0.3125, dt=54.6e-6, nsamps=1 << 20, nbits=4,
            mean=8.0, sigma=2.5, dms=("generate", 0.0, 1000.0, 40.0, 1.25))
CFG3 = dict(nchans=1024, f0=1530.0, df=-0.3125, dt=54.6e-6, nsamps=1 << 20, nbits=8,
            mean=128.0, sigma=20.0, dms=("linear", 0.0, 0.5, 2048))
