"""Writes the committed golden fixtures of tests/golden/*.npz.

The expected outputs here are computed by numpy implementations written for
this script, independently of oracle/ (which is a C restatement): the tests
then check that the oracle and the GPU path both reproduce these vectors.

  plan_tables.npz  delay tables (float32 bits) and DM lists of the five
                   workload geometries, computed in numpy float32/float64 with
                   the operation order of src/Plan.cpp:45-70, 109-119.
  tdd_small.npz    tiny TDD cases for every in_nbits x out_nbits, including a
                   partial killmask: inputs, DM lists and expected outputs
                   (src/tdd/dedisperse/dedisperse_kernel.cuh:45-198).
  fdd_small.npz    tiny 8-bit FDD cases: expected float64 time series and
                   Fourier-domain output (src/fdd/FDDGPUPlan.cpp:85-93, 394-407,
                   618-636; fdd_kernel.cuh:226-229), via numpy.fft.

Run:  python tests/golden/gen_golden.py
"""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
f32 = np.float32

GEOMS = {  # name: (nchans, dt, f0, df)
    "cfg1": (256, 64e-6, 1581.8, -0.390625),
    "cfg2": (1024, 54.6e-6, 1530.0, -0.3125),
    "cfg4": (512, 81.92e-6, 1500.0, -0.78125),
    "cfg5": (1024, 64e-6, 1530.0, -0.3125),
}


def np_delay_table(nchans, dt, f0, df):
    dt, f0, df = f32(dt), f32(f0), -abs(f32(df))
    c = np.arange(nchans, dtype=np.uint64).astype(f32)
    a = (f32(1) / (f0 + c * df)).astype(f32)
    b = f32(1) / f0
    diff = (a * a - b * b).astype(f32)
    k = 4.15e3 / np.float64(dt)
    return (k * diff.astype(np.float64)).astype(f32)


def np_dm_list(dm_start, dm_end, dt, ti, f0, df, nchans, tol):
    dt = float(f32(dt)) * 1e6
    f0, df = float(f32(f0)), float(f32(df))
    f = (f0 + ((nchans // 2) - 0.5) * df) * 1e-3
    tol2 = tol * tol
    a = 8.3 * df / (f * f * f)
    a2 = a * a
    b2 = a2 * (nchans * nchans / 16.0)
    c = (dt * dt + ti * ti) * (tol2 - 1.0)
    out = [f32(dm_start)]
    while out[-1] < f32(dm_end):
        prev = float(out[-1])
        k = c + tol2 * a2 * prev * prev
        dm = (b2 * prev + math.sqrt(-a2 * b2 * prev * prev + (a2 + b2) * k)) / (a2 + b2)
        if not f32(dm) > out[-1]:
            break
        out.append(f32(dm))
    return np.array(out, f32)


def np_max_delay(dm_last, delay_last):
    return int(float(f32(dm_last) * f32(delay_last)) + 0.5)


def np_tdd_delays(dms, delays):
    p = (np.asarray(dms, f32)[:, None] * np.asarray(delays, f32)[None, :]).astype(f32)
    return np.where(p > 0, np.rint(p), 0).astype(np.int64)


def np_tdd(samples, delays, killmask, dms, maxd, in_nbits, out_nbits):
    """samples uint32 [nsamps][nchans]. Brute-force sums, then the epilogue in float32."""
    nsamps, nchans = samples.shape
    nsc = nsamps - maxd
    d = np_tdd_delays(dms, delays)
    km = np.asarray(killmask, np.int64)
    t = np.arange(nsc)[:, None]
    cols = np.arange(nchans)[None, :]
    sums = []
    for i in range(len(dms)):
        idx = t + d[i][None, :]
        v = np.where(idx < nsamps, samples[np.minimum(idx, nsamps - 1), cols], 0).astype(np.uint64)
        terms = (km.astype(np.uint64)[None, :] * v) & 0xFFFFFFFF
        s = terms.sum(axis=1, dtype=np.uint64)
        if in_nbits != 32:
            s &= 0xFFFFFFFF
        sums.append(s)
    sums = np.array(sums)
    in_range = f32(0xFFFFFFFF if in_nbits == 32 else (1 << in_nbits) - 1)
    out_range = f32(1.0) if out_nbits == 32 else f32((1 << out_nbits) - 1)
    factor = (f32(3.0) * f32(1024.0)) / f32(255.0) / f32(16.0)
    den = (in_range * f32(nchans)).astype(f32)
    fs = sums.astype(f32)
    num = (fs * out_range).astype(f32)
    q = (num / den).astype(f32)
    scaled = (q * factor).astype(f32)
    if out_nbits == 32:
        return scaled
    v = np.minimum(np.maximum(scaled.astype(np.float64), 0.0), float(out_range)).astype(f32)
    return np.trunc(v).astype(np.uint8 if out_nbits == 8 else np.uint16)


def np_pack(samples, nbits):
    nsamps, nchans = samples.shape
    if nbits >= 8:
        return samples.astype({8: "<u1", 16: "<u2", 32: "<u4"}[nbits]).view(np.uint8).reshape(nsamps, -1)
    per = 8 // nbits
    out = np.zeros((nsamps, nchans // per), np.uint8)
    for k in range(per):
        out |= (samples[:, k::per].astype(np.uint8) << (k * nbits)).astype(np.uint8)
    return out


def np_fdd(samples_u8, delays, dms, maxd):
    """(time series float64 [ndm][nsamps-maxd], fourier float64 [ndm][N])."""
    nsamps, nchans = samples_u8.shape
    x = (samples_u8.astype(np.float64) - 127.5) / nchans          # [t][c]
    d = np.trunc((np.asarray(dms, f32)[:, None] * np.asarray(delays, f32)[None, :]).astype(f32)).astype(np.int64)
    # time series: exact shift-sum through the Fourier shift theorem at N >= nsamps + max delay
    Nts = nsamps + int(d.max()) + 1
    Xts = np.fft.rfft(x, n=Nts, axis=0)                            # [k][c]
    k = np.arange(Xts.shape[0])[:, None]
    ts = []
    for i in range(len(dms)):
        Y = (Xts * np.exp(2j * np.pi * k * d[i][None, :] / Nts)).sum(axis=1)
        ts.append(np.fft.irfft(Y, n=Nts)[: nsamps - maxd])
    # Fourier output of the fork: N = round_up(nsamps + maxd, 16384), bins 0..N/2-1
    N = ((nsamps + maxd + 16383) // 16384) * 16384
    X = np.fft.rfft(x, n=N, axis=0)
    kk = np.arange(X.shape[0])[:, None]
    four = []
    for i in range(len(dms)):
        Y = (X * np.exp(2j * np.pi * ((kk * d[i][None, :]) % N) / N)).sum(axis=1)[: N // 2]
        row = np.empty(N)
        row[0::2] = Y.real
        row[1::2] = Y.imag
        four.append(row)
    return np.array(ts), np.array(four)


def main():
    rng = np.random.default_rng(20260501)

    # ---- plan tables ----
    tabs = {}
    for name, (nch, dt, f0, df) in GEOMS.items():
        tabs[f"{name}_delay_bits"] = np_delay_table(nch, dt, f0, df).view(np.uint32)
    nch, dt, f0, df = GEOMS["cfg2"]
    dml = np_dm_list(0.0, 1000.0, dt, 40.0, f0, df, nch, 1.25)
    tabs["cfg2_dm_list_bits"] = dml.view(np.uint32)
    tabs["cfg2_max_delay"] = np.array([np_max_delay(dml[-1], np_delay_table(*GEOMS["cfg2"])[-1])])
    np.savez_compressed(HERE / "plan_tables.npz", **tabs)

    # ---- TDD small cases ----
    tdd = {}
    cases = []
    nch, dt, f0, df = 64, 64e-6, 1581.8, -6.25  # wide band: delays of tens of samples
    delays = np_delay_table(nch, dt, f0, df)
    for in_nbits in (1, 2, 4, 8, 16, 32):
        for out_nbits in (8, 16, 32):
            for km_kind in ("ones", "partial"):
                nsamps = 400
                hi = (1 << in_nbits) - 1 if in_nbits < 32 else 0xFFFFFFFF
                if in_nbits == 32:
                    smp = rng.integers(0, 1 << 31, size=(nsamps, nch), dtype=np.uint64).astype(np.uint32) * 2
                else:
                    smp = rng.integers(0, hi + 1, size=(nsamps, nch)).astype(np.uint32)
                dms = np.array([0.0, 0.7, 3.0, 5.5, 8.0], f32)
                km = np.ones(nch, np.int32)
                if km_kind == "partial":
                    km[rng.choice(nch, 11, replace=False)] = 0
                maxd = np_max_delay(dms[-1], delays[-1])
                out = np_tdd(smp, delays, km, dms, maxd, in_nbits, out_nbits)
                key = f"n{in_nbits}_o{out_nbits}_{km_kind}"
                tdd[key + "_in"] = np_pack(smp, in_nbits)
                tdd[key + "_dms"] = dms
                tdd[key + "_km"] = km
                tdd[key + "_out"] = out
                cases.append(key)
    tdd["geometry"] = np.array([nch, dt, f0, df], np.float64)
    tdd["cases"] = np.array(cases)
    np.savez_compressed(HERE / "tdd_small.npz", **tdd)

    # ---- FDD small cases ----
    fdd = {}
    nch, dt, f0, df = 64, 64e-6, 1581.8, -6.25
    delays = np_delay_table(nch, dt, f0, df)
    nsamps = 2000
    smp = np.clip(np.rint(rng.normal(128, 20, size=(nsamps, nch))), 0, 255).astype(np.uint8)
    dms = np.array([0.0, 4.0, 9.9], f32)
    maxd = np_max_delay(dms[-1], delays[-1])
    ts, four = np_fdd(smp, delays, dms, maxd)
    fdd.update(geometry=np.array([nch, dt, f0, df], np.float64), input=smp, dms=dms,
               max_delay=np.array([maxd]), timeseries=ts, fourier=four)
    np.savez_compressed(HERE / "fdd_small.npz", **fdd)
    print("wrote", [p.name for p in HERE.glob("*.npz")])


if __name__ == "__main__":
    main()
