"""The oracle: CPU restatement of the reference dedispersion algorithms.

TEST INFRASTRUCTURE ONLY. Imported by tests/, by __graft_entry__.smoke() and by
the cpu_baseline leg of bench.py, always as the checker; the product
(dedisp_tests_amd, libdedisp.so) never imports or links it.

C parts (dedisp_oracle.c, fft_cpu.c, fdd_cpu.c) are built into
libdedisp_oracle.so by oracle/Makefile; the numpy parts below restate the
reference's Fourier-domain quantities in float64. Each function names the
reference file:line it follows; tests/test_oracle.py pins every one of them
(golden values, closed forms, textbook identities, brute force on tiny cases).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

_DIR = Path(__file__).resolve().parent
LIB_PATH = _DIR / "libdedisp_oracle.so"
_lib = None

_f32p = C.POINTER(C.c_float)


def build() -> Path:
    """Compiles libdedisp_oracle.so with oracle/Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", str(_DIR), "-B", "libdedisp_oracle.so"], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        l = C.CDLL(str(LIB_PATH))
        l.oracle_delay_table.argtypes = [C.c_ulong, C.c_float, C.c_float, C.c_float, C.c_void_p]
        l.oracle_delay_table.restype = None
        l.oracle_generate_dm_list.argtypes = [C.c_float, C.c_float, C.c_double, C.c_double, C.c_double,
                                              C.c_double, C.c_ulong, C.c_double, C.c_void_p, C.c_ulong]
        l.oracle_generate_dm_list.restype = C.c_ulong
        l.oracle_max_delay.argtypes = [C.c_float, C.c_float]
        l.oracle_max_delay.restype = C.c_ulong
        l.oracle_tdd_delay.argtypes = [C.c_float, C.c_float]
        l.oracle_tdd_delay.restype = C.c_uint
        l.oracle_fdd_delay.argtypes = [C.c_float, C.c_float]
        l.oracle_fdd_delay.restype = C.c_int
        l.oracle_tdd.argtypes = [C.c_ulong, C.c_void_p, C.c_void_p, C.c_void_p, C.c_ulong, C.c_ulong,
                                 C.c_ulong, C.c_ulong, C.c_void_p, C.c_uint, C.c_ulong, C.c_void_p,
                                 C.c_uint, C.c_ulong]
        l.oracle_tdd.restype = C.c_int
        l.oracle_fdd_timeseries.argtypes = [C.c_ulong, C.c_void_p, C.c_void_p, C.c_ulong, C.c_ulong,
                                            C.c_ulong, C.c_void_p, C.c_double, C.c_void_p]
        l.oracle_fdd_timeseries.restype = C.c_int
        l.oracle_fdd_circular.argtypes = [C.c_ulong, C.c_void_p, C.c_void_p, C.c_ulong, C.c_ulong,
                                          C.c_ulong, C.c_void_p, C.c_double, C.c_void_p]
        l.oracle_fdd_circular.restype = C.c_int
        l.oracle_fdd_cpu_asis.argtypes = [C.c_ulong, C.c_float, C.c_void_p, C.c_void_p, C.c_ulong,
                                          C.c_ulong, C.c_ulong, C.c_ulong, C.c_void_p, C.c_void_p,
                                          C.c_void_p]
        l.oracle_fdd_cpu_asis.restype = C.c_int
        l.oracle_fdd_cpu_fft_backend.argtypes = []
        l.oracle_fdd_cpu_fft_backend.restype = C.c_char_p
        for suf, ct in (("d", C.c_double), ("f", C.c_float)):
            for name in (f"oracle_rfft_{suf}", f"oracle_irfft_{suf}"):
                getattr(l, name).argtypes = [C.c_ulong, C.c_void_p, C.c_void_p]
                getattr(l, name).restype = None
            getattr(l, f"oracle_cfft_{suf}").argtypes = [C.c_ulong, C.c_void_p, C.c_int]
            getattr(l, f"oracle_cfft_{suf}").restype = None
        _lib = l
    return _lib


def _p(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


# ---------------------------------------------------------------------------
# Plan core (src/Plan.cpp)
# ---------------------------------------------------------------------------

def delay_table(nchans: int, dt: float, f0: float, df: float) -> np.ndarray:
    """src/Plan.cpp:16-40, 109-119 (C restatement)."""
    out = np.empty(nchans, np.float32)
    lib().oracle_delay_table(nchans, dt, f0, df, _p(out))
    return out


def dm_list(dm_start, dm_end, dt, ti, f0, df, nchans, tol) -> np.ndarray:
    """src/Plan.cpp:45-70 as called from :136-154 (dt, f0, df as float32 values)."""
    n = lib().oracle_generate_dm_list(dm_start, dm_end, float(np.float32(dt)), ti,
                                      float(np.float32(f0)), float(np.float32(df)), nchans, tol, None, 0)
    out = np.empty(n, np.float32)
    lib().oracle_generate_dm_list(dm_start, dm_end, float(np.float32(dt)), ti,
                                  float(np.float32(f0)), float(np.float32(df)), nchans, tol, _p(out), n)
    return out


def max_delay(dm_last: float, delay_last: float) -> int:
    """src/Plan.cpp:85-86."""
    return int(lib().oracle_max_delay(dm_last, delay_last))


def tdd_delays(dms: np.ndarray, delays: np.ndarray) -> np.ndarray:
    """__float2uint_rn(dm*delay[c]) of src/tdd/dedisperse/dedisperse_kernel.cuh:163-165, [ndm][nchans]."""
    l = lib()
    return np.array([[l.oracle_tdd_delay(float(d), float(x)) for x in delays] for d in dms], np.int64)


def fdd_delays(dms: np.ndarray, delays: np.ndarray) -> np.ndarray:
    """(int)(dm*delay[c]) of src/fdd/dedisperse/fdd_kernel.cuh:226-229, [ndm][nchans]."""
    p = (np.asarray(dms, np.float32)[:, None] * np.asarray(delays, np.float32)[None, :]).astype(np.float32)
    return np.trunc(p).astype(np.int64)


# ---------------------------------------------------------------------------
# TDD
# ---------------------------------------------------------------------------

def tdd(nchans, delays, killmask, dms, nsamps, max_delay_, data, in_nbits, out_nbits=32,
        first_dm=0, dm_count=None, in_stride=None) -> np.ndarray:
    """Brute-force TDD (src/tdd/TDDPlan.cpp:148-487, dedisperse_kernel.cuh:45-198).
    data: uint8 time-major packed rows. Returns [dm_count][nsamps-max_delay] of the
    output type (uint8 / uint16 / float32)."""
    dms = np.ascontiguousarray(dms, np.float32)
    delays = np.ascontiguousarray(delays, np.float32)
    km = np.ascontiguousarray(killmask if killmask is not None else np.ones(nchans), np.int32)
    if dm_count is None:
        dm_count = dms.size - first_dm
    if in_stride is None:
        in_stride = nchans * in_nbits // 8
    data = np.ascontiguousarray(data, np.uint8)
    nsc = nsamps - max_delay_
    dt = {8: np.uint8, 16: np.uint16, 32: np.float32}[out_nbits]
    out = np.zeros((dm_count, nsc), dt)
    rc = lib().oracle_tdd(nchans, _p(delays), _p(km), _p(dms), first_dm, dm_count, nsamps, max_delay_,
                          _p(data), in_nbits, in_stride, _p(out), out_nbits, out.strides[0])
    assert rc == 0
    return out


# ---------------------------------------------------------------------------
# FDD
# ---------------------------------------------------------------------------

def fdd_timeseries(nchans, delays, dms, nsamps, max_delay_, data_u8, nbits: int = 8) -> np.ndarray:
    """FP64 FDD time series (src/fdd/FDDGPUPlan.cpp:618-636 time-series mode,
    fdd_kernel.cuh:226-229 truncated delays): float64 [ndm][nsamps-max_delay].
    data_u8: one sample value per byte [nsamps][nchans]; samples enter as
    (u - (2^nbits - 1)/2)/nchans (127.5 for the reference's 8-bit input)."""
    dms = np.ascontiguousarray(dms, np.float32)
    delays = np.ascontiguousarray(delays, np.float32)
    data_u8 = np.ascontiguousarray(data_u8, np.uint8)
    out = np.zeros((dms.size, nsamps - max_delay_), np.float64)
    rc = lib().oracle_fdd_timeseries(nchans, _p(delays), _p(dms), dms.size, nsamps, max_delay_,
                                     _p(data_u8), ((1 << nbits) - 1) / 2.0, _p(out))
    assert rc == 0
    return out


def fdd_circular(nchans, delays, dms, nsamps, N, data_u8, nbits: int = 8) -> np.ndarray:
    """FP64 circular shift-sum over length N: float64 [ndm][N] (offset as in
    fdd_timeseries)."""
    dms = np.ascontiguousarray(dms, np.float32)
    delays = np.ascontiguousarray(delays, np.float32)
    data_u8 = np.ascontiguousarray(data_u8, np.uint8)
    out = np.zeros((dms.size, N), np.float64)
    rc = lib().oracle_fdd_circular(nchans, _p(delays), _p(dms), dms.size, nsamps, N, _p(data_u8),
                                   ((1 << nbits) - 1) / 2.0, _p(out))
    assert rc == 0
    return out


def fdd_fourier_length(nsamps: int, max_delay_: int) -> int:
    """N = round_up(nsamp + max_delay, 16384) (src/fdd/FDDGPUPlan.cpp:85-93)."""
    return ((nsamps + max_delay_ + 16383) // 16384) * 16384


def fdd_fourier(nchans, delays, dms, nsamps, max_delay_, data_u8) -> np.ndarray:
    """FP64 Fourier-domain output of the fork's FDD GPU path (src/fdd/FDDGPUPlan.cpp:85-93,
    394-407, 618-650): per DM, N float values = interleaved (re, im) of bins 0..N/2-1 of
    sum_c X_c[k] e^{+2 pi i k d_c/N}, i.e. the DFT of the circular shift-sum. float64 [ndm][N]."""
    N = fdd_fourier_length(nsamps, max_delay_)
    s = fdd_circular(nchans, delays, dms, nsamps, N, data_u8)
    Y = np.fft.rfft(s, axis=1)[:, : N // 2]
    out = np.empty((dms.size, N), np.float64)
    out[:, 0::2] = Y.real
    out[:, 1::2] = Y.imag
    return out


def fdd_cpu_asis(nchans, dt, delays, dms, nsamps, max_delay_, data_u8, dm_first=0, dm_count=None,
                 times: np.ndarray | None = None) -> np.ndarray:
    """FDDCPUPlan::execute_cpu as written (src/fdd/FDDCPUPlan.cpp:439-624), float32
    [dm_count][nsamps-max_delay] (see fdd_cpu.c). times (float64[3]) receives the
    seconds spent in forward transforms, contraction, inverse transforms."""
    dms = np.ascontiguousarray(dms, np.float32)
    delays = np.ascontiguousarray(delays, np.float32)
    data_u8 = np.ascontiguousarray(data_u8, np.uint8)
    if dm_count is None:
        dm_count = dms.size - dm_first
    out = np.zeros((dm_count, nsamps - max_delay_), np.float32)
    rc = lib().oracle_fdd_cpu_asis(nchans, dt, _p(delays), _p(dms), dm_first, dm_count, nsamps, max_delay_,
                                   _p(data_u8), _p(out), None if times is None else _p(times))
    assert rc == 0
    return out


def fdd_cpu_fft_backend() -> str:
    """Which transforms fdd_cpu_asis runs: "fftw3f" (the host's libfftw3f.so.3,
    as the reference links) or "own" (fft_cpu.c)."""
    return lib().oracle_fdd_cpu_fft_backend().decode()


def fdd_cpu_asis_fp64(nchans, dt, delays, dms, nsamps, max_delay_, data_u8) -> np.ndarray:
    """FP64 evaluation of the CPU plan's own formula (src/fdd/FDDCPUPlan.cpp:439-624):
    x_c = (u8-127.5)/nchan zero-padded to N = round_up(nsamp+1,16384); phasors
    exp(i 2 pi f_k tdm_c) with f_k = k/(nsamp dt), tdm_c = dm*delay[c]*dt on bins
    k <= nsamp/2 only (others zero); inverse length-N real transform / N."""
    N = ((nsamps + 1 + 16383) // 16384) * 16384
    nfreq = nsamps // 2 + 1
    x = (data_u8.astype(np.float64) - 127.5) / nchans      # [nsamps][nchans]
    X = np.fft.rfft(x, n=N, axis=0)[:nfreq]                 # [nfreq][nchans]
    k = np.arange(nfreq, dtype=np.float64)
    f = k / (nsamps * dt)
    out = np.empty((len(dms), nsamps - max_delay_))
    for i, dm in enumerate(dms):
        tdm = float(dm) * delays.astype(np.float64) * dt
        Y = (X * np.exp(2j * np.pi * f[:, None] * tdm[None, :])).sum(axis=1)
        full = np.zeros(N // 2 + 1, np.complex128)
        full[:nfreq] = Y
        out[i] = np.fft.irfft(full, n=N)[: nsamps - max_delay_]
    return out


def rfft(x: np.ndarray, double: bool = True) -> np.ndarray:
    """oracle_rfft_{d,f}: real transform of even length n -> n/2+1 complex."""
    n = x.size
    dt = np.float64 if double else np.float32
    xi = np.ascontiguousarray(x, dt)
    out = np.empty(2 * (n // 2 + 1), dt)
    getattr(lib(), "oracle_rfft_d" if double else "oracle_rfft_f")(n, _p(xi), _p(out))
    return out[0::2] + 1j * out[1::2]


def irfft(X: np.ndarray, n: int, double: bool = True) -> np.ndarray:
    """oracle_irfft_{d,f}: n/2+1 complex -> n real, unnormalized (n times the inverse)."""
    dt = np.float64 if double else np.float32
    buf = np.empty(2 * (n // 2 + 1), dt)
    buf[0::2] = X.real
    buf[1::2] = X.imag
    out = np.empty(n, dt)
    getattr(lib(), "oracle_irfft_d" if double else "oracle_irfft_f")(n, _p(buf), _p(out))
    return out


def cfft(z: np.ndarray, sign: int = -1) -> np.ndarray:
    """oracle_cfft_d: complex transform, e^{sign 2 pi i k t / n}, unnormalized."""
    buf = np.empty(2 * z.size, np.float64)
    buf[0::2] = z.real
    buf[1::2] = z.imag
    lib().oracle_cfft_d(z.size, _p(buf), sign)
    return buf[0::2] + 1j * buf[1::2]
