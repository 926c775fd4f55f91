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


# Workload geometries: 256 channels x 65536 8-bit samples; 1024 channels x 2^20
# samples, 4-bit (generated DM list) and 8-bit (2048 linear DM trials).
CFG1 = dict(nchans=256, f0=1581.8, df=-0.390625, dt=64e-6, nsamps=65536, nbits=8,
            mean=128.0, sigma=20.0, dms=("linear", 0.0, 1.0, 256))
CFG2 = dict(nchans=1024, f0=1530.0, df=-0.3125, dt=54.6e-6, nsamps=1 << 20, nbits=4,
            mean=8.0, sigma=2.5, dms=("generate", 0.0, 1000.0, 40.0, 1.25))
CFG3 = dict(nchans=1024, f0=1530.0, df=-0.3125, dt=54.6e-6, nsamps=1 << 20, nbits=8,
            mean=128.0, sigma=20.0, dms=("linear", 0.0, 0.5, 2048))
# Many beams of a smaller telescope, 2-bit samples (nsamps log-uniform in
# 2^17..2^20 per beam); 1024 linear trials from 0 to 2000 (step 2000/1023).
CFG4 = dict(nchans=512, f0=1500.0, df=-0.78125, dt=81.92e-6, nbits=2, mean=1.5, sigma=0.8,
            beams=64, nsamps_range=(1 << 17, 1 << 20), dms=("linspace", 0.0, 2000.0, 1024))
# Streaming: 32 beams delivering 1 s chunks (15625 samples at 64 us), 8-bit,
# 3000 linear trials from 0 to 1500.
CFG5 = dict(nchans=1024, f0=1530.0, df=-0.3125, dt=64e-6, nbits=8, mean=128.0, sigma=20.0,
            beams=32, chunk=15625, dms=("linspace", 0.0, 1500.0, 3000))
