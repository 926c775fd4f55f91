"""Pins the oracle before it is trusted as the checker.

Each oracle function is checked against something other than its own text:
the committed golden vectors (computed independently in numpy by
tests/golden/gen_golden.py), figures quoted from the reference's workload
geometry, closed forms and textbook identities, and brute force on tiny cases.
"""
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).resolve().parent / "golden"
f32 = np.float32


# ---------------------------------------------------------------------------
# Plan core
# ---------------------------------------------------------------------------

GEOMS = {
    "cfg1": (256, 64e-6, 1581.8, -0.390625),
    "cfg2": (1024, 54.6e-6, 1530.0, -0.3125),
    "cfg4": (512, 81.92e-6, 1500.0, -0.78125),
    "cfg5": (1024, 64e-6, 1530.0, -0.3125),
}


@pytest.mark.parametrize("name", list(GEOMS))
def test_delay_table_golden_bits(orc, name):
    g = np.load(GOLD / "plan_tables.npz")
    nch, dt, f0, df = GEOMS[name]
    got = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    assert np.array_equal(got.view(np.uint32), g[f"{name}_delay_bits"])


@pytest.mark.parametrize("name", list(GEOMS))
def test_delay_table_physics(orc, name):
    nch, dt, f0, df = GEOMS[name]
    d = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    assert d[0] == 0.0
    assert np.all(np.diff(d) > 0)                      # df < 0: later channels lag more
    f = float(f32(f0)) - np.arange(nch) * abs(float(f32(df)))
    exact = 4.15e3 / float(f32(dt)) * (1.0 / f**2 - 1.0 / float(f32(f0)) ** 2)
    # float32 a*a - b*b cancels: absolute error bounded by a few ulp of 1/f0^2 scaled
    tol = 4.15e3 / float(f32(dt)) * 4 * np.finfo(np.float32).eps / float(f32(f0)) ** 2 + 1e-6 * exact
    assert np.all(np.abs(d - exact) <= tol)


def test_delay_table_sign_of_df_ignored(orc):
    a = orc.delay_table(256, f32(64e-6), f32(1581.8), f32(-0.390625))
    b = orc.delay_table(256, f32(64e-6), f32(1581.8), f32(0.390625))
    assert np.array_equal(a, b)                        # src/Plan.cpp:31 forces df < 0


def test_reference_geometry_figures(orc):
    # Figures quoted for the workload geometries (delay of the last channel,
    # DM count and max delay of the generated list).
    d1 = orc.delay_table(256, f32(64e-6), f32(1581.8), f32(-0.390625))
    assert abs(d1[255] - 3.600349) < 1e-6
    d2 = orc.delay_table(1024, f32(54.6e-6), f32(1530.0), f32(-0.3125))
    assert abs(d2[1023] - 19.417946) < 1e-5
    dml = orc.dm_list(0.0, 1000.0, 54.6e-6, 40.0, 1530.0, -0.3125, 1024, 1.25)
    assert dml.size == 1167
    assert abs(dml[-1] - 1002.41) < 0.01
    assert orc.max_delay(dml[-1], d2[-1]) == 19465
    assert orc.max_delay(255.0, d1[-1]) == 918
    assert orc.max_delay(1023.5, d2[-1]) == 19874
    d4 = orc.delay_table(512, f32(81.92e-6), f32(1500.0), f32(-0.78125))
    assert orc.max_delay(2000.0, d4[-1]) == 38585
    d5 = orc.delay_table(1024, f32(64e-6), f32(1530.0), f32(-0.3125))
    assert orc.max_delay(1500.0, d5[-1]) == 24849


def test_dm_list_golden_and_invariants(orc):
    g = np.load(GOLD / "plan_tables.npz")
    dml = orc.dm_list(0.0, 1000.0, 54.6e-6, 40.0, 1530.0, -0.3125, 1024, 1.25)
    assert np.array_equal(dml.view(np.uint32), g["cfg2_dm_list_bits"])
    assert dml[0] == 0.0
    assert np.all(np.diff(dml) > 0)
    assert dml[-2] < 1000.0 <= dml[-1]
    assert int(g["cfg2_max_delay"][0]) == 19465


def test_dm_list_tolerance_spacing(orc):
    # Larger tolerance -> coarser list; the list always brackets dm_end.
    a = orc.dm_list(0.0, 500.0, 64e-6, 40.0, 1581.8, -0.390625, 256, 1.1)
    b = orc.dm_list(0.0, 500.0, 64e-6, 40.0, 1581.8, -0.390625, 256, 1.5)
    assert a.size > b.size > 2
    for l in (a, b):
        assert l[-2] < 500.0 <= l[-1]


def test_tdd_delay_rounding(orc):
    # round half to even of the float product, saturating at 0
    assert orc.lib().oracle_tdd_delay(1.0, 2.5) == 2
    assert orc.lib().oracle_tdd_delay(1.0, 3.5) == 4
    assert orc.lib().oracle_tdd_delay(1.0, 3.4999998) == 3
    assert orc.lib().oracle_tdd_delay(-1.0, 3.0) == 0
    assert orc.lib().oracle_tdd_delay(0.0, 7.0) == 0
    # FDD truncates instead (fdd_kernel.cuh:229)
    assert orc.lib().oracle_fdd_delay(1.0, 3.99) == 3


# ---------------------------------------------------------------------------
# TDD brute force
# ---------------------------------------------------------------------------

def _tdd_fixture_cases():
    g = np.load(GOLD / "tdd_small.npz")
    return [str(c) for c in g["cases"]]


@pytest.mark.parametrize("case", _tdd_fixture_cases())
def test_tdd_golden(orc, case):
    g = np.load(GOLD / "tdd_small.npz")
    nch, dt, f0, df = g["geometry"]
    nch = int(nch)
    in_nbits = int(case.split("_")[0][1:])
    out_nbits = int(case.split("_")[1][1:])
    delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    dms = g[case + "_dms"]
    data = g[case + "_in"]
    nsamps = data.shape[0]
    maxd = orc.max_delay(dms[-1], delays[-1])
    out = orc.tdd(nch, delays, g[case + "_km"], dms, nsamps, maxd, data, in_nbits, out_nbits)
    assert np.array_equal(out, g[case + "_out"])


def _pack(samples, nbits):
    import dedisp_tests_amd as dd   # pure-numpy helper, no library load
    return dd.pack_samples(samples, nbits)


def test_tdd_dm0_is_channel_sum(orc):
    rng = np.random.default_rng(1)
    nch, nsamps = 96, 300
    s = rng.integers(0, 16, size=(nsamps, nch)).astype(np.uint32)
    delays = orc.delay_table(nch, f32(64e-6), f32(1581.8), f32(-1.0))
    out = orc.tdd(nch, delays, None, np.array([0.0], f32), nsamps, 0, _pack(s, 4), 4, 32)
    sums = s.sum(axis=1).astype(f32)
    expect = (sums * f32(1.0) / (f32(15.0) * f32(nch))) * ((f32(3.0) * f32(1024.0)) / f32(255.0) / f32(16.0))
    assert np.array_equal(out[0], expect.astype(f32))


def test_tdd_all_15_gives_factor(orc):
    # 4-bit all-15 input, f32 output: E = 15*nchans/(15*nchans) * factor = factor
    nch, nsamps = 1024, 64
    s = np.full((nsamps, nch), 15, np.uint32)
    delays = orc.delay_table(nch, f32(54.6e-6), f32(1530.0), f32(-0.3125))
    out = orc.tdd(nch, delays, None, np.array([0.0, 1.0], f32), nsamps, orc.max_delay(1.0, delays[-1]),
                  _pack(s, 4), 4, 32)
    factor = (f32(3.0) * f32(1024.0)) / f32(255.0) / f32(16.0)
    assert np.all(out == factor)


def test_tdd_killmask_zero_and_constant(orc):
    nch, nsamps = 64, 200
    delays = orc.delay_table(nch, f32(64e-6), f32(1581.8), f32(-2.0))
    dms = np.array([0.0, 2.0, 4.0], f32)
    maxd = orc.max_delay(dms[-1], delays[-1])
    s = np.full((nsamps, nch), 200, np.uint32)
    out = orc.tdd(nch, delays, np.zeros(nch, np.int32), dms, nsamps, maxd, _pack(s, 8), 8, 8)
    assert np.all(out == 0)
    out = orc.tdd(nch, delays, None, dms, nsamps, maxd, _pack(s, 8), 8, 16)
    # constant v: sum = nchans*v -> E = (nch*200)*65535/(255*nch)*factor, clipped, truncated
    factor = (f32(3.0) * f32(1024.0)) / f32(255.0) / f32(16.0)
    v = (f32(nch * 200) * f32(65535.0) / (f32(255.0) * f32(nch))) * factor
    assert np.all(out == np.uint16(np.trunc(min(float(v), 65535.0))))


def test_tdd_impulse_lands_at_delay(orc):
    nch, nsamps = 64, 400
    delays = orc.delay_table(nch, f32(64e-6), f32(1581.8), f32(-6.25))
    dm = f32(7.3)
    dms = np.array([dm], f32)
    maxd = orc.max_delay(dm, delays[-1])
    c, t0 = 50, 250
    s = np.zeros((nsamps, nch), np.uint32)
    s[t0, c] = 255
    out = orc.tdd(nch, delays, None, dms, nsamps, maxd, _pack(s, 8), 8, 32)
    d = int(np.rint(np.float32(dm * delays[c])))
    nz = np.nonzero(out[0])[0]
    assert list(nz) == [t0 - d]


# ---------------------------------------------------------------------------
# FDD
# ---------------------------------------------------------------------------

def test_fdd_timeseries_golden(orc):
    g = np.load(GOLD / "fdd_small.npz")
    nch, dt, f0, df = g["geometry"]
    nch = int(nch)
    delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    data = g["input"]
    maxd = int(g["max_delay"][0])
    ts = orc.fdd_timeseries(nch, delays, g["dms"], data.shape[0], maxd, data)
    assert np.max(np.abs(ts - g["timeseries"])) < 1e-9


def test_fdd_fourier_golden(orc):
    g = np.load(GOLD / "fdd_small.npz")
    nch, dt, f0, df = g["geometry"]
    nch = int(nch)
    delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    data = g["input"]
    maxd = int(g["max_delay"][0])
    four = orc.fdd_fourier(nch, delays, g["dms"], data.shape[0], maxd, data)
    ref = g["fourier"]
    assert np.linalg.norm(four - ref) / np.linalg.norm(ref) < 1e-12


def test_fdd_circular_shift_identity(orc):
    # DFT of the circular shift-sum equals sum_c X_c[k] e^{+2 pi i k d_c / N}
    rng = np.random.default_rng(3)
    nch, nsamps, N = 8, 100, 128
    data = rng.integers(0, 256, size=(nsamps, nch)).astype(np.uint8)
    delays = orc.delay_table(nch, f32(64e-6), f32(1400.0), f32(-20.0))
    dms = np.array([0.0, 3.3], f32)
    s = orc.fdd_circular(nch, delays, dms, nsamps, N, data)
    x = (data.astype(np.float64) - 127.5) / nch
    X = np.fft.fft(x, n=N, axis=0)
    d = orc.fdd_delays(dms, delays)
    k = np.arange(N)[:, None]
    for i in range(len(dms)):
        Y = (X * np.exp(2j * np.pi * k * d[i][None, :] / N)).sum(axis=1)
        assert np.allclose(np.fft.fft(s[i]), Y, atol=1e-9)


@pytest.mark.parametrize("n", [8, 30, 1056, 9240, 16384, 81920])
def test_fft_against_numpy(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    X = orc.rfft(x)
    assert np.allclose(X, np.fft.rfft(x), atol=1e-9 * np.sqrt(n))
    back = orc.irfft(X, n) / n
    assert np.allclose(back, x, atol=1e-10)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.allclose(orc.cfft(z, -1), np.fft.fft(z), atol=1e-9 * np.sqrt(n))


def test_fft_impulse_and_parseval(orc):
    n = 2 * 3 * 5 * 7 * 11 * 4
    x = np.zeros(n)
    x[0] = 1.0
    assert np.allclose(orc.rfft(x), 1.0)
    rng = np.random.default_rng(5)
    y = rng.standard_normal(n)
    Y = np.fft.fft(y)
    Yo = orc.cfft(y.astype(np.complex128), -1)
    assert abs(np.sum(np.abs(Yo) ** 2) / n - np.sum(y**2)) < 1e-8 * np.sum(y**2)
    assert np.allclose(Yo, Y, atol=1e-8)


def test_fdd_cpu_asis_matches_its_fp64_formula(orc):
    rng = np.random.default_rng(7)
    nch, nsamps = 64, 6000
    data = np.clip(np.rint(rng.normal(128, 20, size=(nsamps, nch))), 0, 255).astype(np.uint8)
    dt = f32(64e-6)
    delays = orc.delay_table(nch, dt, f32(1581.8), f32(-1.5))
    dms = np.array([0.0, 5.0, 12.5], f32)
    maxd = orc.max_delay(dms[-1], delays[-1])
    got = orc.fdd_cpu_asis(nch, dt, delays, dms, nsamps, maxd, data)
    ref = orc.fdd_cpu_asis_fp64(nch, float(dt), delays, dms, nsamps, maxd, data)
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert rel < 1e-4, rel
    # It is not the dedispersed series: only bins 0..nsamp/2 of a longer
    # transform are filled, which low-passes the output even at DM 0.
    ts = orc.fdd_timeseries(nch, delays, dms, nsamps, maxd, data)
    assert np.linalg.norm(got[0] - ts[0]) / np.linalg.norm(ts[0]) > 0.1
