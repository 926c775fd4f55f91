"""Benchmark of the dedispersion hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Headline workload (the FDD precision form): 1024 channels from 1530 MHz down in
0.3125 MHz steps, dt = 54.6 us, 2^20 8-bit samples per channel (N(128, 20)
with five dispersed pulses), 2048 linear DM trials 0..1023.5, overlap-save
segments of 2^16 samples, time-series output. A step is one dedisp_execute of
the FDD plan with the bfloat16 matrix-core contraction (hi/lo split phasors
and spectra, three products, float32 accumulation), input and output resident
in HBM (DEDISP_DEVICE_POINTERS): transpose, forward transforms, contraction,
inverse transforms. Value: DM-trials x output samples per second. Parity: the
rel-L2 error against the FP64 oracle on 8 spread DM rows must stay within
1e-3 and the S/N of the five pulses within 0.5 %; the bench exits non-zero
otherwise.

Prints one JSON line. Besides the contract fields it carries:
  roofline       matrix-core view of the contraction kernel: algorithmic flops
                 (8 per complex multiply-add) / its time measured with HIP
                 events on the plan's stream; issued_frac counts the three
                 products and the padded chunk tiles the kernel issues
  fdd            the float32 and float16 contraction forms on the same
                 workload, each with its error against the FP64 oracle and its
                 roofline, and the reference fork's default Fourier output
  tdd            TDD on 1024 ch x 2^20 4-bit samples, the 1167 trials of
                 dedisp_generate_dm_list(0, 1000, 40 us, 1.25): bit for bit
                 against the oracle on 10 rows, HBM and integer-add rooflines
                 of the channel-sum kernel, and its host-pointer run
  batch          64 beams of uneven length through one TDD graph replay, and
                 FDD-BF16 on all 64 with their chunks concatenated (one graph),
                 with its error against the FP64 oracle
  stream         the streaming front end on 32 beams x 1 s chunks, 1024 ch,
                 3000 DMs: push latency percentiles and real-time factor, per
                 beam (TDD) and with one batched push per round (TDD,
                 FDD-BF16, FDD-FP32)
  end_to_end     the headline execute with host pointers (the reference's
                 calling convention): PCIe-inclusive rate and stage times
  filterbank     the dedisp_fil tool on the TDD input written as a SIGPROC
                 file: read, dedisperse 64 trials, write .dat/.inf
  measured_peaks achievable HBM copy bandwidth and bf16 / fp32 GEMM rates on
                 this GPU; the rooflines also give fractions of these
  cpu_baseline   the reference's CPU FDD plan (restated as written, OpenMP)
                 timed on this host on a bounded DM sample of the headline
                 DM list, scaled to all 2048 trials
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0                      # MI355X HBM3E
TRAFFIC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                               "tdd-1024ch-2p20x4bit-1167dm_hbm-traffic.md")


def _profile_value(key: str):
    """An integer `key: value` line of the committed channel-sum counter summary
    (rocprofv3 --pmc runs of this bench), or None."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.strip().startswith(key + ":"):
                    return int(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def measured_traffic():
    """HBM bytes per launch of the channel-sum kernel (FETCH_SIZE / WRITE_SIZE,
    copy-calibrated), or None."""
    return _profile_value("traffic_bytes_per_launch")


def measured_valu_insts():
    """VALU wave-instructions per launch of the channel-sum kernel (SQ_INSTS_VALU), or None."""
    return _profile_value("valu_insts_per_launch")


# VALU issue peak: a full-rate 32-bit VALU op issues every 2 cycles per SIMD on
# gfx950 (measured with two or more waves per SIMD): 1024 SIMDs x 2.4 GHz / 2.
VALU_ISSUE_PEAK = 1024 * 2.4e9 / 2

NCH, DT, F0, DF = 1024, 54.6e-6, 1530.0, -0.3125
NSAMPS = 1 << 20
NBITS = 4


def make_input(torch, seed: int = 1234):
    """Packed 4-bit input [NSAMPS, NCH/2] on the GPU (torch is only plumbing here)."""
    import dedisp_tests_amd as dd
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    x = torch.randn((NSAMPS, NCH), generator=g, device="cuda", dtype=torch.float32)
    s = torch.clamp(torch.round(x * 2.5 + 8.0), 0, 15).to(torch.int16)
    del x
    delays = torch.from_numpy(dd.generate_delay_table(NCH, DT, F0, DF)).cuda()
    rng = np.random.default_rng(seed)
    for _ in range(5):                     # 5 pulses, DM ~ U(50, 900), width 1..16
        dm = float(rng.uniform(50, 900))
        width = int(rng.integers(1, 17))
        t0 = int(rng.integers(0, NSAMPS // 2))
        d = torch.round((torch.tensor(np.float32(dm), device="cuda") * delays)).to(torch.int64)
        cols = torch.arange(NCH, device="cuda")
        for w in range(width):
            t = t0 + d + w
            ok = t < NSAMPS
            s[t[ok], cols[ok]] = torch.clamp(s[t[ok], cols[ok]] + 4, max=15)
    s = s.to(torch.uint8)
    packed = (s[:, 0::2] | (s[:, 1::2] << 4)).contiguous()   # channel 2b low nibble of byte b
    return s, packed


def cpu_baseline(seconds_budget: float = 30.0):
    """Reference CPU FDD plan restated as written (oracle, OpenMP over DMs) on the
    FDD leg's workload (1024 ch x 2^20 8-bit samples, 2048 trials 0..1023.5), a
    sample of DMs; the DM-independent transforms are timed once and the per-DM
    part is scaled to the full list. The host's FFTW is used when it loads."""
    import oracle
    ncores = os.cpu_count() or 1
    try:
        ncores = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    threads = int(os.environ.get("OMP_NUM_THREADS", ncores))
    delays = oracle.delay_table(NCH, np.float32(DT), np.float32(F0), np.float32(DF))
    # the DM list of the FDD leg: 2048 trials 0 .. 1023.5
    dms = (np.arange(2048, dtype=np.float32) * np.float32(0.5)).astype(np.float32)
    maxd = oracle.max_delay(dms[-1], delays[-1])
    nsc = NSAMPS - maxd
    rng = np.random.default_rng(5)
    data = np.clip(np.rint(rng.normal(128, 20, size=(NSAMPS, NCH)).astype(np.float32)), 0, 255).astype(np.uint8)
    sample = max(1, min(threads, 16))
    idx = np.linspace(0, len(dms) - 1, sample).astype(int)
    times = np.zeros(3)
    t0 = time.perf_counter()
    oracle.fdd_cpu_asis(NCH, np.float32(DT), delays, dms[idx], NSAMPS, maxd, data, times=times)
    wall = time.perf_counter() - t0
    per_dm = (times[1] + times[2]) / sample
    full = times[0] + per_dm * len(dms)
    backend = oracle.fdd_cpu_fft_backend()
    fft_note = ("the host's FFTW (libfftw3f.so.3)" if backend == "fftw3f" else
                "own mixed-radix FFT (libfftw3f.so.3 not found on this host)")
    return {
        "value": len(dms) * nsc / full,
        "unit": "DM*samples/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"FDDCPUPlan::execute_cpu restated (oracle/fdd_cpu.c), 1024 ch x 2^20 8-bit "
                   f"samples, {sample} of {len(dms)} DMs timed ({wall:.1f} s wall): forward "
                   f"transforms {times[0]:.2f} s once + {per_dm:.2f} s per DM, scaled to "
                   f"the {len(dms)} DMs of the FDD leg (0..1023.5); transforms: {fft_note}"),
        "fft": backend,
    }


MFMA_PEAK_T = {"fp32": 157.3, "bf16": 2500.0, "fp16": 2500.0}   # dense TFLOP/s


def mfma16_tiles(nchunk: int) -> tuple:
    """N-tiles per workgroup and chunk groups of the 16 x 16 tile contraction
    for nchunk chunks (the kernel's rule: fewest tiles plus 1.5 tiles' worth
    of phasor work per group, at most 3 tiles)."""
    best, best_ntl = None, 1
    for ntl in (1, 2, 3):
        groups = -(-nchunk // (8 * ntl))
        cost = groups * ntl + 1.5 * groups
        if best is None or cost < best or (cost == best and ntl > best_ntl):
            best, best_ntl = cost, ntl
    return best_ntl, -(-nchunk // (8 * best_ntl))


def fdd_leg(torch, dd, steps: int, warmup: int, forms=("fp32", "bf16", "fp16"), e2e: bool = True) -> dict:
    """FDD on 1024 ch x 2^20 8-bit samples (N(128,20), five pulses), 2048 linear DM
    trials 0..1023.5, overlap-save chunks of 2^16 samples, time-series output,
    input and output in HBM. One line per contraction form with its error against
    the FP64 oracle on a sample of DM rows."""
    import oracle
    g = torch.Generator(device="cuda")
    g.manual_seed(99)
    x = torch.randn((NSAMPS, NCH), generator=g, device="cuda", dtype=torch.float32)
    s = torch.clamp(torch.round(x * 20.0 + 128.0), 0, 255)
    del x
    delays_np = dd.generate_delay_table(NCH, DT, F0, DF)
    delays = torch.from_numpy(delays_np).cuda()
    rng = np.random.default_rng(99)
    cols = torch.arange(NCH, device="cuda")
    pulse_dms = []
    for _ in range(5):
        dm = float(rng.uniform(50, 900))
        pulse_dms.append(dm)
        width = int(rng.integers(1, 17))
        t0 = int(rng.integers(0, NSAMPS // 2))
        d = torch.round(torch.tensor(np.float32(dm), device="cuda") * delays).to(torch.int64)
        for w in range(width):
            t = t0 + d + w
            ok = t < NSAMPS
            s[t[ok], cols[ok]] = torch.clamp(s[t[ok], cols[ok]] + 30, max=255)
    d_in = s.to(torch.uint8).contiguous()
    del s
    dms = (np.arange(2048, dtype=np.float32) * np.float32(0.5)).astype(np.float32)
    host_in = d_in.cpu().numpy()
    sample_rows = np.linspace(0, 2047, 8).astype(int).tolist()
    otab = oracle.delay_table(NCH, np.float32(DT), np.float32(F0), np.float32(DF))
    ref = oracle.fdd_timeseries(NCH, otab, dms[sample_rows], NSAMPS, 0, host_in)
    # the DM trials nearest the injected pulses, for the S/N comparison
    pulse_rows = [int(round(dm / 0.5)) for dm in pulse_dms]
    ref_p = oracle.fdd_timeseries(NCH, otab, dms[pulse_rows], NSAMPS, 0, host_in)

    def snr(x):
        return (x.max() - np.median(x)) / x.std()
    out = {"workload": "fdd-1024ch-2^20x8bit-2048dm-nfft65536", "ndm": 2048, "nfft": 65536}
    for name, prec in (("fp32", dd.FDD_FP32), ("bf16", dd.FDD_BF16), ("fp16", dd.FDD_FP16)):
        if name not in forms:
            continue
        plan = dd.Plan(NCH, DT, F0, DF, impl=dd.FDD)
        plan.fdd_set_output(dd.FDD_TIMESERIES)
        plan.fdd_set_segmented(True, 65536)
        plan.fdd_set_precision(prec)
        plan.set_dm_list(dms)
        maxd = plan.max_delay
        nsc = NSAMPS - maxd
        d_out = torch.empty((2048, nsc), dtype=torch.float32, device="cuda")
        plan.set_profiling(True)
        torch.cuda.synchronize()
        for _ in range(warmup):
            assert plan.execute_raw(NSAMPS, d_in, 8, d_out, 32, dd.DEVICE_POINTERS) == 0
        torch.cuda.synchronize()
        prof = []
        t0 = time.perf_counter()
        for _ in range(steps):
            assert plan.execute_raw(NSAMPS, d_in, 8, d_out, 32, dd.DEVICE_POINTERS) == 0
            prof.append(plan.profile())
        torch.cuda.synchronize()
        el = (time.perf_counter() - t0) / steps
        got = d_out[sample_rows].cpu().numpy().astype(np.float64)
        r = ref[:, :nsc]
        rel = float(np.linalg.norm(got - r) / np.linalg.norm(r))
        maxabs = float(np.max(np.abs(got - r)))
        got_p = d_out[pulse_rows].cpu().numpy().astype(np.float64)
        snr_err = max(abs(snr(got_p[j]) / snr(ref_p[j, :nsc]) - 1.0) for j in range(len(pulse_rows)))
        c_ms = float(np.mean([p["dedisperse_ms"] for p in prof]))
        st_ms = {"transpose": float(np.mean([p["unpack_ms"] for p in prof])),
                 "forward_fft": float(np.mean([p["fft_ms"] for p in prof])),
                 "contraction": c_ms,
                 "inverse_fft": float(np.mean([p["ifft_ms"] for p in prof]))}
        # algorithmic HBM bytes of the transform stages: packed input once, the
        # spectra X[c][chunk][bin] and Y[dm][chunk][bin] (complex float), the
        # float output rows
        nchunk = -(-nsc // (65536 - maxd))
        x_bytes = NCH * nchunk * (65536 // 2 + 1) * 8
        y_bytes = 2048 * nchunk * (65536 // 2 + 1) * 8
        st_bytes = {"transpose": 2 * NSAMPS * NCH, "forward_fft": NSAMPS * NCH + x_bytes,
                    "inverse_fft": y_bytes + 2048 * nsc * 4}
        cmacs = prof[-1]["work_units"]
        tflops = 8 * cmacs / (c_ms * 1e-3) / 1e12
        out[name] = {
            "value": 2048 * nsc / el, "ms_per_step": 1e3 * el, "steps": steps,
            "max_delay": maxd, "nchunk": nchunk, "rows_checked": sample_rows, "pulse_rows": pulse_rows,
            "stages_ms": st_ms,
            "stages_hbm": {k: {"algorithmic_bytes": v, "gbs": v / (st_ms[k] * 1e-3) / 1e9}
                           for k, v in st_bytes.items()},
            "rel_l2_vs_fp64": rel, "max_abs_vs_fp64": maxabs,
            "pulse_snr_rel_err": float(snr_err),
            "pulse_snr_note": f"max over the 5 injected pulses of |S/N / S/N(FP64) - 1| on the nearest DM trial "
                              f"(rows {pulse_rows}); S/N = (max - median) / std of the row",
            "roofline": {"bound": "mfma", "achieved": tflops, "peak": MFMA_PEAK_T[name], "unit": "TFLOP/s",
                         "frac": tflops / MFMA_PEAK_T[name], "kernel": "contraction", "kernel_ms": c_ms,
                         "algorithmic_flops": 8 * cmacs,
                         "note": "algorithmic flops: 8 per complex multiply-add of DMs x channels x chunks x "
                                 "bins 0..nfft/2 (no padding, no compensation products); the kernel time is "
                                 "measured with HIP events on the plan's stream over the timed executes"},
        }
        if name == "bf16":
            # matrix-core work the kernel issues: three products per complex
            # multiply-add (hi x hi, hi x lo, lo x hi) over chunk groups padded
            # to 8-chunk N-tiles
            ntl, ngroups = mfma16_tiles(nchunk)
            issued = 8 * cmacs * 3 * (8 * ntl * ngroups) / nchunk
            out[name]["roofline"]["issued_flops"] = issued
            out[name]["roofline"]["issued_frac"] = issued / (c_ms * 1e-3) / 1e12 / MFMA_PEAK_T[name]
        if name == "bf16" and e2e:
            # end-to-end scope: the same execute with the caller's pageable host buffers
            h_out = np.zeros((2048, nsc), np.float32)
            assert plan.execute_raw(NSAMPS, host_in, 8, h_out, 32, 0) == 0
            ts = []
            for _ in range(2):
                t1 = time.perf_counter()
                assert plan.execute_raw(NSAMPS, host_in, 8, h_out, 32, 0) == 0
                ts.append(time.perf_counter() - t1)
            p = plan.profile()
            e_ms = 1e3 * float(np.median(ts))
            out[name]["end_to_end"] = {
                "value": 2048 * nsc / (e_ms * 1e-3), "ms_per_execute": e_ms,
                "h2d_ms": p["h2d_ms"], "d2h_ms": p["d2h_ms"],
                "pcie_bytes": NSAMPS * NCH + 2048 * nsc * 4,
                "equal_to_device_run": bool(np.array_equal(h_out[sample_rows], d_out[sample_rows].cpu().numpy())),
                "note": "host pointers (pageable numpy buffers): upload through page-locked slots, "
                        "download overlapped with the later DM batches"}
            del h_out
        del d_out, plan
        torch.cuda.empty_cache()
    if "fp32" in forms:
        # the reference fork's default output: the whole-series Fourier
        # coefficients, N = round_up(nsamps + max_delay, 16384) floats per DM
        # (bins 0..N/2-1 as re/im), float32 contraction
        plan = dd.Plan(NCH, DT, F0, DF, impl=dd.FDD)
        plan.set_dm_list(dms)
        N = plan.out_length(NSAMPS)
        nsc = NSAMPS - plan.max_delay
        d_out = torch.empty((2048, N), dtype=torch.float32, device="cuda")
        plan.set_profiling(True)
        assert plan.execute_raw(NSAMPS, d_in, 8, d_out, 32, dd.DEVICE_POINTERS) == 0
        torch.cuda.synchronize()
        nst = max(1, steps // 2)
        prof = []
        t0 = time.perf_counter()
        for _ in range(nst):
            assert plan.execute_raw(NSAMPS, d_in, 8, d_out, 32, dd.DEVICE_POINTERS) == 0
            prof.append(plan.profile())
        torch.cuda.synchronize()
        el = (time.perf_counter() - t0) / nst
        got = d_out[sample_rows].cpu().numpy().astype(np.float64)
        refF = oracle.fdd_fourier(NCH, otab, dms[sample_rows], NSAMPS, plan.max_delay, host_in)
        out["fourier_fp32"] = {
            "value": 2048 * nsc / el, "ms_per_step": 1e3 * el, "N": N,
            "stages_ms": {"transpose": float(np.mean([p["unpack_ms"] for p in prof])),
                          "forward_fft": float(np.mean([p["fft_ms"] for p in prof])),
                          "contraction": float(np.mean([p["dedisperse_ms"] for p in prof])),
                          "inverse_and_row_transforms": float(np.mean([p["ifft_ms"] for p in prof]))},
            "rel_l2_vs_fp64": float(np.linalg.norm(got - refF) / np.linalg.norm(refF)),
            "note": "default output mode of the reference fork (whole-series Fourier coefficients); "
                    "value in DM x (nsamps - max_delay) per second like the other forms"}
        del d_out, plan
        torch.cuda.empty_cache()
    if "fp32" in out and "fp16" in out:
        out["speedup_fp16_vs_fp32"] = out["fp16"]["value"] / out["fp32"]["value"]
    return out


def batch_leg(torch, dd, steps: int = 2) -> dict:
    """Batched beams: 64 beams of log-uniform length 2^17..2^20 (seeded), 512
    channels from 1500 MHz down in 0.78125 MHz steps, dt 81.92 us, 2-bit
    samples clip(round(N(1.5, 0.8)), 0, 3), 1024 linear DM trials 0..2000. TDD
    with the autotuned tiling: one transpose per beam and one channel-sum
    launch for all beams, replayed as one HIP graph (device pointers, 8-bit
    output); then FDD-BF16 on all 64 beams (the overlap-save chunks of the
    beams concatenated into forward transform, contraction and inverse
    transform launches, replayed as one HIP graph; float output, 134 GB).
    Inputs and outputs stay in HBM."""
    nb, nch, dt, f0, df = 64, 512, 81.92e-6, 1500.0, -0.78125
    rng = np.random.default_rng(4)
    lengths = [int(2 ** rng.uniform(17, 20)) for _ in range(nb)]
    g = torch.Generator(device="cuda")
    g.manual_seed(4)
    d_in = []
    for n in lengths:
        x = torch.clamp(torch.round(torch.randn((n, nch), generator=g, device="cuda") * 0.8 + 1.5), 0, 3).to(torch.uint8)
        d_in.append((x[:, 0::4] | (x[:, 1::4] << 2) | (x[:, 2::4] << 4) | (x[:, 3::4] << 6)).contiguous())
        del x
    plan = dd.Plan(nch, dt, f0, df, impl=dd.TDD)
    dms = np.linspace(0.0, 2000.0, 1024, dtype=np.float32)
    plan.set_dm_list(dms)
    md = plan.max_delay
    tune = plan.autotune(1 << 20, 2, 8)
    d_out = [torch.empty((1024, n - md), dtype=torch.uint8, device="cuda") for n in lengths]
    torch.cuda.synchronize()

    def step():
        rc = plan.execute_batch_raw(lengths, d_in, 2, d_out, 8, dd.DEVICE_POINTERS)
        if rc != 0:
            raise dd.DedispError(rc, "dedisp_amd_execute_batch")

    step()  # captures the graph
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / steps
    dm_samples = 1024 * sum(n - md for n in lengths)
    # parity spot check: the shortest beam equals its single execute
    i = int(np.argmin(lengths))
    one = torch.empty_like(d_out[i])
    assert plan.execute_raw(lengths[i], d_in[i], 2, one, 8, dd.DEVICE_POINTERS) == 0
    torch.cuda.synchronize()
    parity = "ok" if torch.equal(one, d_out[i]) else "MISMATCH"
    out = {"workload": "batch-64beam-512ch-2^17..2^20x2bit-1024dm", "beams": nb,
           "total_samples": int(sum(lengths)), "max_delay": md, "tdd_tiling": tune,
           "tdd": {"value": dm_samples / el, "unit": "DM*samples/s", "ms_per_batch": 1e3 * el,
                   "launch": "one HIP graph: 64 transposes + 1 channel-sum launch", "out_nbits": 8,
                   "parity": parity}}
    del d_out, one, plan
    torch.cuda.empty_cache()

    # FDD-BF16 on all 64 beams: the overlap-save chunks of every beam go
    # through one forward transform, contraction and inverse transform per
    # group of beams, captured as one HIP graph; error against the oracle on
    # sampled rows of the shortest and the longest beam
    import oracle
    plan = dd.Plan(nch, dt, f0, df, impl=dd.FDD)
    plan.fdd_set_output(dd.FDD_TIMESERIES)
    plan.fdd_set_precision(dd.FDD_BF16)
    plan.set_dm_list(dms)
    ftune = plan.autotune(1 << 20, 2, 32)
    f_out = [torch.empty((1024, n - md), dtype=torch.float32, device="cuda") for n in lengths]
    rc = plan.execute_batch_raw(lengths, d_in, 2, f_out, 32, dd.DEVICE_POINTERS)
    assert rc == 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        assert plan.execute_batch_raw(lengths, d_in, 2, f_out, 32, dd.DEVICE_POINTERS) == 0
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / steps
    otab = oracle.delay_table(nch, np.float32(dt), np.float32(f0), np.float32(df))
    rows = [0, 511, 1023]
    errs = {}
    for b in (int(np.argmin(lengths)), int(np.argmax(lengths))):
        host = dd.unpack_samples(d_in[b].cpu().numpy(), 2, nch)
        ref = oracle.fdd_timeseries(nch, otab, dms[rows], lengths[b], md, host, nbits=2)
        got = f_out[b][rows].cpu().numpy().astype(np.float64)
        errs[f"beam{b}_{lengths[b]}"] = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        del host
    out["fdd_bf16"] = {"value": 1024 * sum(n - md for n in lengths) / el, "unit": "DM*samples/s",
                       "ms_per_batch": 1e3 * el, "nfft": ftune["nfft"], "beams": nb,
                       "launch": "one HIP graph: 64 transposes + per group of beams 1 forward, and per DM batch "
                                 "1 contraction and 1 inverse launch",
                       "rel_l2_vs_fp64": errs,
                       "rel_l2_note": "rows 0, 511, 1023 of the shortest and the longest beam against the oracle"}
    del f_out, plan, d_in
    torch.cuda.empty_cache()
    return out


def stream_leg(torch, dd, rounds: int = 4) -> dict:
    """Streaming front end: 32 beams delivering 1 s chunks (15625 samples at
    64 us) of 1024 channels from 1530 MHz down in 0.3125 MHz steps, 8-bit
    N(128, 20), 3000 linear DM trials 0..1500, TDD plan, float
    output. Round 0 fills the histories and is not timed; every later round
    pushes one chunk per beam with input and output in HBM. Then the same
    workload with one batched push per round: TDD, FDD-BF16 and FDD-FP32."""
    nb, chunk, nch, dt = 32, 15625, 1024, 64e-6
    plan = dd.Plan(nch, dt, 1530.0, -0.3125, impl=dd.TDD)
    plan.set_dm_list(np.linspace(0.0, 1500.0, 3000, dtype=np.float32))
    st = dd.Stream(plan, nb, chunk, 8, 32)
    d_out = torch.empty((plan.dm_count, chunk), dtype=torch.float32, device="cuda")
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    round_s, push_ms = [], []
    for r in range(rounds):
        x = torch.randn((nb, chunk, nch), generator=g, device="cuda", dtype=torch.float32)
        chunks = torch.clamp(torch.round(x * 20.0 + 128.0), 0, 255).to(torch.uint8)
        del x
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for b in range(nb):
            tp = time.perf_counter()
            rc = st.push_raw(b, chunks[b], d_out, dd.DEVICE_POINTERS)  # returns when the push is done
            if rc != 0:
                raise dd.DedispError(rc, "dedisp_amd_stream_push")
            if r > 0:
                push_ms.append(1e3 * (time.perf_counter() - tp))
        if r > 0:
            round_s.append(time.perf_counter() - t0)
        del chunks
    stats = st.stats()
    st.close()
    t_round = float(np.median(round_s))
    lat = np.sort(np.array(push_ms))
    out = {
        "workload": "stream-32beam-1024ch-15625x8bit-3000dm", "beams": nb, "chunk_nsamps": chunk,
        "max_delay": plan.max_delay, "ndm": plan.dm_count, "timed_rounds": len(round_s),
        "round_ms": 1e3 * t_round,
        "real_time_factor": nb * chunk * dt / t_round,
        "dm_samples_per_s": plan.dm_count * chunk * nb / t_round,
        "push_ms": {"p50": _nearest(lat, 0.5), "p99": _nearest(lat, 0.99), "mean": float(lat.mean()),
                    "max": float(lat[-1]), "pushes": len(lat),
                    "note": "timed rounds, host-synchronous pushes, nearest rank"},
        "library_stats_all_rounds": {"p50_ms": stats["p50_ms"], "p99_ms": stats["p99_ms"],
                                     "pushes": stats["pushes"]},
    }
    del d_out, st
    torch.cuda.empty_cache()
    out["tdd_batched"] = stream_batched(torch, dd, plan, rounds=6)
    del plan
    torch.cuda.empty_cache()
    for form, prec in (("fdd_bf16_batched", dd.FDD_BF16), ("fdd_fp32_batched", dd.FDD_FP32)):
        fplan = dd.Plan(nch, dt, 1530.0, -0.3125, impl=dd.FDD)
        fplan.fdd_set_precision(prec)
        fplan.set_dm_list(np.linspace(0.0, 1500.0, 3000, dtype=np.float32))
        out[form] = stream_batched(torch, dd, fplan, rounds=4)
        del fplan
        torch.cuda.empty_cache()
    return out


def _nearest(sorted_vals, q: float) -> float:
    """Nearest-rank percentile of a sorted array."""
    n = len(sorted_vals)
    return float(sorted_vals[min(n - 1, max(0, int(np.ceil(q * n)) - 1))])


def stream_batched(torch, dd, plan, rounds: int) -> dict:
    """The streaming workload with one batched push per round: the 32 beams'
    1 s chunks of a round go through one dedisp_amd_stream_push_batch (one
    channel-sum launch for TDD; for FDD the overlap-save chunks of all beams in
    one forward transform / contraction / inverse sequence). Round 0 fills the
    histories and is not timed; a chunk's latency is its round's call time."""
    nb, chunk, nch, dt = 32, 15625, 1024, 64e-6
    st = dd.Stream(plan, nb, chunk, 8, 32)
    d_out = [torch.empty((plan.dm_count, chunk), dtype=torch.float32, device="cuda") for _ in range(nb)]
    g = torch.Generator(device="cuda")
    g.manual_seed(6)
    call_ms = []
    for r in range(rounds):
        x = torch.randn((nb, chunk, nch), generator=g, device="cuda", dtype=torch.float32)
        chunks = torch.clamp(torch.round(x * 20.0 + 128.0), 0, 255).to(torch.uint8)
        del x
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rc = st.push_batch_raw(list(range(nb)), [chunks[b] for b in range(nb)], d_out, dd.DEVICE_POINTERS)
        if rc != 0:
            raise dd.DedispError(rc, "dedisp_amd_stream_push_batch")
        if r > 0:
            call_ms.append(1e3 * (time.perf_counter() - t0))
        del chunks
    st.close()
    lat = np.sort(np.array(call_ms))
    t_round = float(np.median(lat)) * 1e-3
    return {"timed_rounds": len(lat), "round_ms": 1e3 * t_round,
            "real_time_factor": nb * chunk * dt / t_round,
            "dm_samples_per_s": plan.dm_count * chunk * nb / t_round,
            "chunk_latency_ms": {"p50": _nearest(lat, 0.5), "p99": _nearest(lat, 0.99), "max": float(lat[-1]),
                                 "note": "one batched push per round; every chunk of a round completes with it"}}


def _sigproc_kw(name: str) -> bytes:
    import struct
    return struct.pack("<i", len(name)) + name.encode()


def fil_leg(host_in) -> dict:
    """The filterbank front end end to end: the headline input (1024 ch x 2^20
    4-bit samples) written as a SIGPROC file, then the dedisp_fil tool reads
    it, dedisperses 64 linear DM trials 0..945 with the TDD plan (host
    pointers) and writes one PRESTO .dat/.inf pair per trial. Stage times are
    the tool's own; the files go to a temporary directory that is removed."""
    import shutil
    import struct
    import subprocess
    import tempfile
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dedisp_tests_amd", "bin", "dedisp_fil")
    tmp = tempfile.mkdtemp(prefix="dedisp_fil_")
    try:
        h = _sigproc_kw("HEADER_START") + _sigproc_kw("source_name") + _sigproc_kw("SYNTH")
        for k, v in (("telescope_id", 4), ("machine_id", 10)):
            h += _sigproc_kw(k) + struct.pack("<i", v)
        for k, v in (("tstart", 60000.0), ("tsamp", DT), ("fch1", F0), ("foff", DF)):
            h += _sigproc_kw(k) + struct.pack("<d", v)
        for k, v in (("nchans", NCH), ("nbits", NBITS), ("nifs", 1)):
            h += _sigproc_kw(k) + struct.pack("<i", v)
        h += _sigproc_kw("HEADER_END")
        fil = os.path.join(tmp, "obs.fil")
        with open(fil, "wb") as f:
            f.write(h)
            f.write(np.ascontiguousarray(host_in).tobytes())
        t0 = time.perf_counter()
        r = subprocess.run([exe, "-f", fil, "-o", os.path.join(tmp, "obs"), "-i", "tdd", "-r", "0", "-s", "15",
                            "-n", "64"], capture_output=True, text=True, timeout=600)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            raise RuntimeError(f"dedisp_fil failed: {r.stderr[-500:]}")
        took = {k: float(v) for k, v in re.findall(r"(Reading input file|Dedispersion|Writing DM trials) took ([0-9.]+)",
                                                   r.stdout)}
        m = re.search(r"Computing (\d+) out of", r.stdout)
        nsc = int(m.group(1))
        ndat = len([n for n in os.listdir(tmp) if n.endswith(".dat")])
        out_bytes = 64 * (nsc & ~1) * 4
        return {"workload": "fil-1024ch-2^20x4bit-64dm", "file_bytes": len(h) + host_in.nbytes, "dat_files": ndat,
                "read_s": took.get("Reading input file"), "dedisperse_s": took.get("Dedispersion"),
                "write_s": took.get("Writing DM trials"), "tool_wall_s": wall,
                "dedisperse_dm_samples_per_s": 64 * nsc / took["Dedispersion"],
                "file_to_files_dm_samples_per_s": 64 * nsc / wall,
                "read_gbs": (len(h) + host_in.nbytes) / took["Reading input file"] / 1e9,
                "write_gbs": out_bytes / took["Writing DM trials"] / 1e9,
                "note": "tool wall time includes process start, device set-up and plan creation"}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def e2e_leg(dd, plan, host_in, d_out, ndm: int, nsc: int, steps: int = 3) -> dict:
    """End-to-end scope of the headline workload: the same execute with host
    pointers, the reference's calling convention (pageable numpy buffers):
    upload through page-locked slots, transpose, channel sum in DM blocks whose
    download overlaps the later blocks. PCIe-inclusive; never the headline value."""
    h_out = np.zeros((ndm, nsc), np.float32)
    rc = plan.execute_raw(NSAMPS, host_in, NBITS, h_out, 32, 0)
    if rc != 0:
        raise dd.DedispError(rc, "dedisp_execute")
    ts = []
    for _ in range(steps):
        t0 = time.perf_counter()
        rc = plan.execute_raw(NSAMPS, host_in, NBITS, h_out, 32, 0)
        ts.append(time.perf_counter() - t0)
        if rc != 0:
            raise dd.DedispError(rc, "dedisp_execute")
    p = plan.profile()
    ms = 1e3 * float(np.median(ts))
    pcie = NSAMPS * NCH * NBITS // 8 + ndm * nsc * 4
    return {"value": ndm * nsc / (ms * 1e-3), "unit": "DM*samples/s", "ms_per_execute": ms,
            "h2d_ms": p["h2d_ms"], "channel_sum_ms": p["dedisperse_ms"], "d2h_ms": p["d2h_ms"],
            "pcie_bytes": pcie, "pcie_gbs": pcie / (ms * 1e-3) / 1e9,
            "equal_to_device_run": bool(np.array_equal(h_out, d_out.cpu().numpy())),
            "note": "host pointers (pageable numpy buffers): H2D + transpose + channel sum + D2H"}


def peaks_leg(torch) -> dict:
    """Achievable peaks on this GPU, beside the nominal ones: HBM by a 1 GiB
    device-to-device copy (read + write bytes), the matrix cores by 8192^3
    bf16, fp16 and fp32 GEMMs (torch.matmul on the vendor library; 2 n^3 flops).
    Timed with events on torch's current stream, median of 10."""
    def timed(f, reps=10):
        f()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            f()
            b.record()
            b.synchronize()
            ts.append(a.elapsed_time(b) * 1e-3)
        return float(np.median(ts))
    n_copy = 1 << 28  # floats = 1 GiB
    x = torch.empty(n_copy, dtype=torch.float32, device="cuda").fill_(1.0)
    y = torch.empty_like(x)
    t = timed(lambda: y.copy_(x))
    out = {"hbm_copy_gbs": 2 * 4 * n_copy / t / 1e9}
    del x, y
    n = 8192
    for name, dt in (("bf16", torch.bfloat16), ("fp16", torch.float16), ("fp32", torch.float32)):
        a = torch.randn((n, n), device="cuda").to(dt)
        b = torch.randn((n, n), device="cuda").to(dt)
        t = timed(lambda: torch.matmul(a, b), reps=5)
        out[f"{name}_gemm_tflops"] = 2.0 * n ** 3 / t / 1e12
        del a, b
    torch.cuda.empty_cache()
    out.update(int_add_peak())
    out["note"] = ("nominal: HBM 8000 GB/s, bf16 / fp16 2500 and fp32 157.3 TFLOP/s dense, "
                   "32-bit integer adds 256 CU x 4 SIMD x 64 lanes / 2 cycles x 2.4 GHz = 7.86e13/s")
    return out


def int_add_peak() -> dict:
    """Achievable 32-bit integer add rate: libdd_peaks.so's probe (8 independent
    v_add_u32 chains per thread, 8 waves on every SIMD), fastest of 5 runs."""
    import ctypes
    lib = ctypes.CDLL(os.path.join(os.path.dirname(os.path.abspath(__file__)), "dedisp_tests_amd", "libdd_peaks.so"))
    f = lib.dd_peak_int_add
    f.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    f.restype = ctypes.c_int
    adds, insts = ctypes.c_double(), ctypes.c_double()
    rc = f(65536, 4096, 5, ctypes.byref(adds), ctypes.byref(insts))
    if rc != 0:
        raise RuntimeError(f"dd_peak_int_add failed: HIP error {rc}")
    return {"int_add_lane_adds_per_s": adds.value, "valu_wave_insts_per_s": insts.value}


def progress(leg: str, t_start: float) -> None:
    """One line on stderr when a leg has finished (stdout carries only the JSON line)."""
    print(f"bench: {leg} leg done at {time.perf_counter() - t_start:.1f} s", file=sys.stderr, flush=True)


def valu_roofline(k_ms: float, adds_t: float) -> dict:
    """VALU issue rate of the channel-sum kernel: the measured instruction count
    per launch (committed counter summary) over the kernel time, against the
    full-rate issue peak. Channel-adds/s are reported beside it (the packed lanes
    add 4 samples per 32-bit lane operation)."""
    n = measured_valu_insts()
    out = {"bound": "valu", "unit": "wave-instructions/s", "peak": VALU_ISSUE_PEAK,
           "adds_per_s": adds_t * 1e12,
           "source": "profiles/tdd-1024ch-2p20x4bit-1167dm_hbm-traffic.md (SQ_INSTS_VALU)"}
    if n is None:
        out.update({"achieved": None, "frac": None})
    else:
        ach = n / (k_ms * 1e-3)
        out.update({"achieved": ach, "frac": ach / VALU_ISSUE_PEAK, "valu_insts_per_launch": n})
    return out


BF16_TOL = 1e-3        # stated tolerance of the bfloat16 form: rel-L2 vs the FP64 oracle
SNR_TOL = 0.005        # injected-pulse S/N preserved to 0.5 %
METRIC = "DM·samples/s: FDD BF16-MFMA vs FDD FP32 vs TDD at stated err; roofline %"
CONTRACTION_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                                   "fdd-bf16-1024ch-2p20x8bit-2048dm_contraction-counters.md")


def contraction_traffic():
    """HBM bytes per launch of the bfloat16 contraction kernel (FETCH_SIZE /
    WRITE_SIZE, copy-calibrated) from the committed counter summary, or None."""
    try:
        with open(CONTRACTION_PROFILE) as f:
            for line in f:
                if line.strip().startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def tdd_leg(torch, dd, steps: int, warmup: int) -> tuple:
    """TDD on 1024 ch x 2^20 4-bit samples, the 1167-trial DM list of
    dedisp_generate_dm_list(0, 1000, 40 us, 1.25), 32-bit output, input and
    output in HBM; parity bit for bit on DM 0 (an independent channel sum) and
    9 spread rows against the oracle. Returns (leg dict, plan, packed input,
    output)."""
    s, d_in = make_input(torch)
    plan = dd.Plan(NCH, DT, F0, DF, impl=dd.TDD)
    plan.generate_dm_list(0.0, 1000.0, 40.0, 1.25)
    ndm, maxd = plan.dm_count, plan.max_delay
    nsc = NSAMPS - maxd
    d_out = torch.empty((ndm, nsc), dtype=torch.float32, device="cuda")
    plan.set_profiling(True)
    torch.cuda.synchronize()

    def step():
        rc = plan.execute_raw(NSAMPS, d_in, NBITS, d_out, 32, dd.DEVICE_POINTERS)
        if rc != 0:
            raise dd.DedispError(rc, "dedisp_execute")

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    kernel_ms, unpack_ms = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
        p = plan.profile()
        kernel_ms.append(p["dedisperse_ms"] / max(1, p["dedisperse_launches"]))
        unpack_ms.append(p["unpack_ms"])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0

    # parity: DM 0 is the plain channel sum (independent torch sum), and 9
    # spread DM rows, the largest DM included, equal the oracle bit for bit
    row0 = d_out[0].cpu().numpy()
    sums = s[:nsc].to(torch.int32).sum(dim=1).cpu().numpy().astype(np.float32)
    f32 = np.float32
    factor = (f32(3.0) * f32(1024.0)) / f32(255.0) / f32(16.0)
    ref0 = ((sums * f32(1.0)) / (f32(15.0) * f32(NCH)) * factor).astype(f32)
    import oracle
    parity_rows = sorted(set(np.linspace(0, ndm - 1, 9).astype(int).tolist()))
    host_packed = d_in.cpu().numpy()
    otab = oracle.delay_table(NCH, np.float32(DT), np.float32(F0), np.float32(DF))
    ref_rows = oracle.tdd(NCH, otab, None, plan.dm_list()[parity_rows], NSAMPS, maxd, host_packed, NBITS, 32)
    got_rows = d_out[parity_rows].cpu().numpy()
    bad = [r for j, r in enumerate(parity_rows) if not np.array_equal(got_rows[j], ref_rows[j])]
    parity = "ok" if (np.array_equal(row0, ref0) and not bad) else "MISMATCH"
    del host_packed, got_rows, ref_rows, s

    k_ms = float(np.mean(kernel_ms))
    in_bytes = NSAMPS * NCH * NBITS / 8
    out_bytes = ndm * nsc * 4
    achieved_gbs = (in_bytes + out_bytes) / (k_ms * 1e-3) / 1e9
    adds_t = ndm * nsc * NCH / (k_ms * 1e-3) / 1e12
    leg = {
        "value": ndm * nsc / (elapsed / steps), "unit": "DM*samples/s", "ms_per_step": 1e3 * elapsed / steps,
        "steps": steps, "dtype": "u32",
        "config": {"workload": "tdd-1024ch-2^20x4bit-1167dm", "nchans": NCH, "nsamps": NSAMPS,
                   "in_nbits": NBITS, "out_nbits": 32, "ndm": ndm, "max_delay": maxd,
                   "dm_list": "dedisp_generate_dm_list(0,1000,40us,1.25)"},
        "roofline": {"bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved_gbs / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "traffic_source": "profiles/tdd-1024ch-2p20x4bit-1167dm_hbm-traffic.md",
                     "kernel": "tdd_runs_k", "kernel_ms": k_ms, "algorithmic_bytes": in_bytes + out_bytes,
                     "note": "LDS reads and VALU issue bound this kernel, not HBM: see valu_roofline"},
        "valu_roofline": valu_roofline(k_ms, adds_t),
        "stages_ms": {"transpose": float(np.mean(unpack_ms)), "channel_sum": k_ms},
        "parity": parity,
        "parity_detail": {"rows_checked": len(parity_rows) + 1, "rows": [0] + parity_rows, "mismatched_rows": bad,
                          "against": "oracle TDD (bit for bit) and an independent channel sum for DM 0"},
    }
    return leg, plan, d_in, d_out, ndm, nsc


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tdd", action="store_true")
    ap.add_argument("--no-fdd", action="store_true",
                    help="run the TDD leg only and print it (profiling the TDD kernels)")
    ap.add_argument("--no-stream", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-e2e", action="store_true",
                    help="skip the host-pointer runs (profiling: they launch the same kernels on DM blocks)")
    ap.add_argument("--fdd-forms", default="fp32,bf16,fp16",
                    help="contraction forms of the FDD leg besides the headline bfloat16 (counter runs use bf16)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy 1 GiB device to device twice (the known-size kernel that "
                         "scales FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the dedispersion path runs on a single device)", file=sys.stderr)
        sys.exit(2)

    import torch
    import dedisp_tests_amd as dd

    if args.calibrate_copy:
        x = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(2):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes": x.numel() * 4, "copies": 2}))
        return

    t_start = time.perf_counter()
    if args.no_fdd:
        tdd = tdd_leg(torch, dd, args.steps, args.warmup)[0]
        print(json.dumps({"tdd": tdd}))
        sys.exit(0 if tdd["parity"] == "ok" else 1)
    # ---- headline: FDD with the bfloat16 matrix-core contraction ----
    forms = tuple(dict.fromkeys(["bf16"] + [f for f in args.fdd_forms.split(",") if f]))
    fdd = fdd_leg(torch, dd, args.steps, args.warmup, forms=forms, e2e=not args.no_e2e)
    b = fdd.pop("bf16")
    parity = "ok" if (b["rel_l2_vs_fp64"] <= BF16_TOL and b["pulse_snr_rel_err"] <= SNR_TOL) else "MISMATCH"
    roof = dict(b["roofline"])
    roof["traffic"] = contraction_traffic()
    roof["traffic_source"] = os.path.relpath(CONTRACTION_PROFILE, os.path.dirname(os.path.abspath(__file__)))
    line = {
        "metric": METRIC,
        "value": b["value"],
        "unit": "DM*samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": b["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic",
        "config": {"workload": fdd["workload"], "nchans": NCH, "nsamps": NSAMPS, "in_nbits": 8,
                   "out_nbits": 32, "ndm": 2048, "dm_list": "linear 0..1023.5 step 0.5", "nfft": 65536,
                   "max_delay": b["max_delay"], "nchunk": b["nchunk"], "output": "time series",
                   "contraction": "bfloat16 MFMA, hi/lo split phasors and spectra (three products), "
                                  "float32 accumulation"},
        "roofline": roof,
        "stages_ms": b["stages_ms"],
        "stages_hbm": b["stages_hbm"],
        "rel_l2_vs_fp64": b["rel_l2_vs_fp64"],
        "max_abs_vs_fp64": b["max_abs_vs_fp64"],
        "pulse_snr_rel_err": b["pulse_snr_rel_err"],
        "tolerance": {"rel_l2": BF16_TOL, "pulse_snr": SNR_TOL},
        "parity": parity,
        "parity_detail": {"rows": b["rows_checked"], "pulse_rows": b["pulse_rows"],
                          "against": "oracle.fdd_timeseries (FP64 delays, phases and transforms)"},
    }
    if "end_to_end" in b:
        line["end_to_end"] = b["end_to_end"]
    line["fdd"] = fdd
    if "fp32" in fdd:
        line["fdd"]["speedup_bf16_vs_fp32"] = b["value"] / fdd["fp32"]["value"]
    progress("fdd", t_start)
    if parity != "ok":
        print(json.dumps(line))
        print(f"bench: FDD-BF16 error {b['rel_l2_vs_fp64']:.3g} / pulse S/N {b['pulse_snr_rel_err']:.3g} "
              "outside the stated tolerance", file=sys.stderr)
        sys.exit(1)
    torch.cuda.empty_cache()
    peaks = peaks_leg(torch)
    line["measured_peaks"] = peaks
    line["roofline"]["frac_of_measured_gemm"] = line["roofline"]["achieved"] / peaks["bf16_gemm_tflops"]
    for form, key in (("fp32", "fp32_gemm_tflops"), ("fp16", "fp16_gemm_tflops")):
        if form in line["fdd"]:
            rl = line["fdd"][form]["roofline"]
            rl["frac_of_measured_gemm"] = rl["achieved"] / peaks[key]
    for st in line["stages_hbm"].values():
        st["frac_of_measured_hbm"] = st["gbs"] / peaks["hbm_copy_gbs"]
    progress("peaks", t_start)

    # ---- TDD beside it (bit-exact) ----
    if not args.no_tdd:
        tdd, plan, d_in, d_out, ndm, nsc = tdd_leg(torch, dd, args.steps, args.warmup)
        tdd["roofline"]["frac_of_measured_peak"] = tdd["roofline"]["achieved"] / peaks["hbm_copy_gbs"]
        vr = tdd["valu_roofline"]
        vr["adds_frac_of_measured_int_add"] = vr["adds_per_s"] / peaks["int_add_lane_adds_per_s"]
        if vr["achieved"] is not None:
            vr["frac_of_measured_issue"] = vr["achieved"] / peaks["valu_wave_insts_per_s"]
        line["tdd"] = tdd
        progress("tdd", t_start)
        if tdd["parity"] != "ok":
            print(json.dumps(line))
            print(f"bench: TDD output differs from the oracle on rows {tdd['parity_detail']['mismatched_rows']}",
                  file=sys.stderr)
            sys.exit(1)
        if not args.no_e2e:
            host_in = d_in.cpu().numpy()
            tdd["end_to_end"] = e2e_leg(dd, plan, host_in, d_out, ndm, nsc)
            progress("end_to_end", t_start)
            line["filterbank"] = fil_leg(host_in)
            progress("filterbank", t_start)
            del host_in
        del d_out, d_in, plan
        torch.cuda.empty_cache()
    if not args.no_batch:
        line["batch"] = batch_leg(torch, dd)
        torch.cuda.empty_cache()
        progress("batch", t_start)
    if not args.no_stream:
        line["stream"] = stream_leg(torch, dd)
        torch.cuda.empty_cache()
        progress("stream", t_start)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline()
        progress("cpu_baseline", t_start)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
