"""FDD parity on the GPU against the FP64 oracle.

Tolerance (the FDD-FP32 bar): relative L2 error <= 1e-4 against the float64
shift-sum; the reference's own float32 phase formula stays about 1e-3 from it,
so this is a tighter bar than bit-equality with the reference would give.
"""
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).resolve().parent / "golden"
f32 = np.float32
pytestmark = pytest.mark.gpu

REL_FP32 = 1e-4


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def _fdd_plan(dd, nch, dt, f0, df, nfft=0, prec=None):
    plan = dd.Plan(nch, dt, f0, df, impl=dd.FDD)
    plan.fdd_set_output(dd.FDD_TIMESERIES)
    if nfft:
        plan.fdd_set_segmented(True, nfft)
    if prec is not None:
        plan.fdd_set_precision(prec)
    return plan


def test_fdd_golden_timeseries(dd):
    g = np.load(GOLD / "fdd_small.npz")
    nch, dt, f0, df = g["geometry"]
    plan = _fdd_plan(dd, int(nch), dt, f0, df)
    plan.set_dm_list(g["dms"])
    assert plan.max_delay == int(g["max_delay"][0])
    out = plan.execute(g["input"], 8, 32)
    assert out.shape == g["timeseries"].shape
    assert rel_l2(out, g["timeseries"]) < REL_FP32


@pytest.mark.parametrize("nfft", [8192, 16384, 32768, 65536])
def test_fdd_segment_lengths_vs_oracle(dd, orc, nfft):
    rng = np.random.default_rng(nfft)
    nch, nsamps = 128, 70000 + 123
    dt, f0, df = 64e-6, 1581.8, -0.78125
    data = np.clip(np.rint(rng.normal(128, 20, size=(nsamps, nch))), 0, 255).astype(np.uint8)
    delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    dms = np.linspace(0, 450, 21, dtype=f32)
    plan = _fdd_plan(dd, nch, dt, f0, df, nfft=nfft)
    plan.set_dm_list(dms)
    out = plan.execute(data, 8, 32)
    ref = orc.fdd_timeseries(nch, delays, dms, nsamps, plan.max_delay, data)
    assert rel_l2(out, ref) < REL_FP32


def test_fdd_cfg1_timeseries_and_pulse(dd, orc):
    from dedisp_tests_amd import synth
    nch, dt, f0, df = 256, 64e-6, 1581.8, -0.390625
    delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    s = synth.filterbank(65536, nch, 8, 128.0, 20.0, seed=1, pulses=[(100.0, 30000, 4, 30)], delays=delays)
    data = s.astype(np.uint8)
    dms = np.arange(256, dtype=f32)
    plan = _fdd_plan(dd, nch, dt, f0, df)
    plan.set_dm_list(dms)
    out = plan.execute(data, 8, 32)
    ref = orc.fdd_timeseries(nch, delays, dms, 65536, plan.max_delay, data)
    assert rel_l2(out, ref) < REL_FP32
    # per-DM max error well below the noise
    assert np.max(np.abs(out - ref)) < 1e-3
    # injected pulse: S/N at DM 100 within 0.5 % of the oracle's
    def snr(x):
        return (x.max() - np.median(x)) / x.std()
    assert abs(snr(out[100]) / snr(ref[100]) - 1) < 5e-3
    assert int(np.argmax(out.max(axis=1))) in (99, 100, 101)


def test_fdd_use_segmented_env_and_killmask(dd, orc, monkeypatch):
    rng = np.random.default_rng(4)
    nch, nsamps = 64, 30000
    dt, f0, df = 64e-6, 1581.8, -2.0
    data = np.clip(np.rint(rng.normal(128, 20, size=(nsamps, nch))), 0, 255).astype(np.uint8)
    delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    dms = np.linspace(0, 60, 11, dtype=f32)
    monkeypatch.setenv("USE_SEGMENTED", "1")
    plan = dd.Plan(nch, dt, f0, df, impl=dd.FDD)      # default output, segmented by env
    plan.set_dm_list(dms)
    assert plan.out_length(nsamps) == nsamps - plan.max_delay
    km = np.ones(nch, np.int32)
    km[[3, 17, 40]] = 0
    plan.set_killmask(km)
    out = plan.execute(data, 8, 32)
    keep = np.nonzero(km)[0]
    ref = orc.fdd_timeseries(len(keep), delays[keep], dms, nsamps, plan.max_delay,
                             np.ascontiguousarray(data[:, keep])) * (len(keep) / nch)
    assert rel_l2(out, ref) < REL_FP32


def test_fdd_device_pointers(dd, orc):
    import torch
    rng = np.random.default_rng(8)
    nch, nsamps = 128, 50000
    dt, f0, df = 64e-6, 1581.8, -0.78125
    data = np.clip(np.rint(rng.normal(128, 20, size=(nsamps, nch))), 0, 255).astype(np.uint8)
    delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    dms = np.linspace(0, 300, 40, dtype=f32)
    plan = _fdd_plan(dd, nch, dt, f0, df)
    plan.set_dm_list(dms)
    nsc = nsamps - plan.max_delay
    d_in = torch.from_numpy(data).cuda()
    d_out = torch.zeros((dms.size, nsc), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    assert plan.execute_raw(nsamps, d_in, 8, d_out, 32, dd.DEVICE_POINTERS) == 0
    ref = orc.fdd_timeseries(nch, delays, dms, nsamps, plan.max_delay, data)
    assert rel_l2(d_out.cpu().numpy(), ref) < REL_FP32


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
@pytest.mark.parametrize("fourier", [False, True])
def test_fdd_outputs_stay_inside_the_buffer(dd, orc, prec, fourier):
    """Device-pointer outputs with guard regions before and after them: the
    chunk inverses (stores bounded by a buffer record count, the last chunk
    of every row partly valid) and the row transforms write nothing outside
    the rows, in every contraction form and both output modes."""
    import torch
    rng = np.random.default_rng(81)
    nch, nsamps = 128, 50000
    dt, f0, df = 64e-6, 1581.8, -0.78125
    data = np.clip(np.rint(rng.normal(128, 20, size=(nsamps, nch))), 0, 255).astype(np.uint8)
    delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    dms = np.linspace(0, 300, 40, dtype=f32)
    plan = dd.Plan(nch, dt, f0, df, impl=dd.FDD)
    plan.fdd_set_output(dd.FDD_FOURIER if fourier else dd.FDD_TIMESERIES)
    plan.fdd_set_precision({"fp32": dd.FDD_FP32, "bf16": dd.FDD_BF16, "fp16": dd.FDD_FP16}[prec])
    plan.set_dm_list(dms)
    n_out = plan.out_length(nsamps)
    guard, sentinel = 8192, 12345.5
    buf = torch.full((2 * guard + dms.size * n_out,), sentinel, dtype=torch.float32, device="cuda")
    d_out = buf[guard:guard + dms.size * n_out].view(dms.size, n_out)
    d_in = torch.from_numpy(data).cuda()
    torch.cuda.synchronize()
    assert plan.execute_raw(nsamps, d_in, 8, d_out, 32, dd.DEVICE_POINTERS) == 0
    torch.cuda.synchronize()
    host = buf.cpu().numpy()
    assert np.all(host[:guard] == sentinel) and np.all(host[guard + dms.size * n_out:] == sentinel)
    out = host[guard:guard + dms.size * n_out].reshape(dms.size, n_out)
    if fourier:
        ref = orc.fdd_fourier(nch, delays, dms, nsamps, plan.max_delay, data)
    else:
        ref = orc.fdd_timeseries(nch, delays, dms, nsamps, plan.max_delay, data)
    assert rel_l2(out, ref) < PREC_TOL[prec]


def test_fdd_validation(dd):
    plan = _fdd_plan(dd, 64, 64e-6, 1581.8, -2.0)
    buf = np.zeros((1000, 64), np.uint8)
    out = np.zeros(100000, np.float32)
    assert plan.execute_raw(1000, buf, 8, out, 32) == dd.DEDISP_NO_DM_LIST_SET
    plan.set_dm_list(np.array([0.0, 5.0], f32))
    assert plan.execute_raw(1000, buf, 16, out, 32) == dd.DEDISP_UNSUPPORTED_IN_NBITS
    assert plan.execute_raw(1000, buf, 8, out, 16) == dd.DEDISP_UNSUPPORTED_OUT_NBITS
    assert plan.execute_raw(1000, None, 8, out, 32) == dd.DEDISP_INVALID_POINTER
    # FDD plans have no gulp / adv / guru interface (src/clib/dedisp.cpp:182, 428, 488)
    assert dd.lib().dedisp_set_gulp_size(plan.handle, 10) == dd.DEDISP_INVALID_PLAN
    assert plan.execute_adv(1000, buf, 8, 64, out, 32, 4000) == dd.DEDISP_INVALID_PLAN


# Tolerances of the contraction forms against the FP64 oracle (relative L2):
# float32 MFMA and the three-product bfloat16 form carry float32-level error;
# the single-product float16 form carries the float16 rounding of phasors and
# spectra (about 2^-11 per term).
PREC_TOL = {"fp32": 1e-4, "bf16": 1e-4, "fp16": 1e-3}


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
def test_fdd_contraction_forms_vs_oracle(dd, orc, prec):
    from dedisp_tests_amd import synth
    p = {"fp32": dd.FDD_FP32, "bf16": dd.FDD_BF16, "fp16": dd.FDD_FP16}[prec]
    nch, dt, f0, df = 256, 64e-6, 1581.8, -0.390625
    delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    s = synth.filterbank(65536, nch, 8, 128.0, 20.0, seed=7, pulses=[(60.0, 20000, 4, 30)], delays=delays)
    data = s.astype(np.uint8)
    dms = np.arange(0, 300, 1.1, dtype=f32)     # 273 DMs: a partial DM block
    plan = _fdd_plan(dd, nch, dt, f0, df, prec=p)
    plan.set_dm_list(dms)
    out = plan.execute(data, 8, 32)
    ref = orc.fdd_timeseries(nch, delays, dms, 65536, plan.max_delay, data)
    err = rel_l2(out, ref)
    assert err < PREC_TOL[prec], err
    i = int(np.argmin(np.abs(dms - 60.0)))
    snr = lambda x: (x.max() - np.median(x)) / x.std()
    assert abs(snr(out[i]) / snr(ref[i]) - 1) < 5e-3


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
def test_fdd_forms_ragged_channels_many_chunks(dd, orc, prec):
    # 100 channels (padded to 112 channel rows in the contraction), 2^13-sample
    # segments over 4e5 samples: 51 chunks, i.e. a full and a partly filled
    # group of 32 chunks; 150 DMs fill part of a DM block
    p = {"fp32": dd.FDD_FP32, "bf16": dd.FDD_BF16, "fp16": dd.FDD_FP16}[prec]
    rng = np.random.default_rng(17)
    nch, nsamps = 100, 400000
    dt, f0, df = 64e-6, 1581.8, -1.0
    data = np.clip(np.rint(rng.normal(128, 20, size=(nsamps, nch))), 0, 255).astype(np.uint8)
    delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    dms = np.linspace(0, 150, 150, dtype=f32)
    plan = _fdd_plan(dd, nch, dt, f0, df, nfft=8192, prec=p)
    plan.set_dm_list(dms)
    assert (nsamps - plan.max_delay + (8192 - plan.max_delay) - 1) // (8192 - plan.max_delay) > 32
    out = plan.execute(data, 8, 32)
    ref = orc.fdd_timeseries(nch, delays, dms, nsamps, plan.max_delay, data)
    assert rel_l2(out, ref) < PREC_TOL[prec]


def test_fdd_valu_and_mfma_fp32_agree(dd, orc, monkeypatch):
    rng = np.random.default_rng(12)
    nch, nsamps = 96, 40000
    dt, f0, df = 64e-6, 1581.8, -1.0
    data = np.clip(np.rint(rng.normal(128, 20, size=(nsamps, nch))), 0, 255).astype(np.uint8)
    dms = np.linspace(0, 100, 37, dtype=f32)
    outs = []
    for valu in ("1", "0"):
        monkeypatch.setenv("DEDISP_FDD_VALU", valu)
        plan = _fdd_plan(dd, nch, dt, f0, df)
        plan.set_dm_list(dms)
        outs.append(plan.execute(data, 8, 32))
    delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    ref = orc.fdd_timeseries(nch, delays, dms, nsamps, plan.max_delay, data)
    assert rel_l2(outs[0], ref) < REL_FP32
    assert rel_l2(outs[1], ref) < REL_FP32


# ---------------------------------------------------------------------------
# Whole-series Fourier-coefficient output (the reference fork's default)
# ---------------------------------------------------------------------------

def test_fdd_fourier_golden(dd):
    g = np.load(GOLD / "fdd_small.npz")
    nch, dt, f0, df = g["geometry"]
    plan = dd.Plan(int(nch), dt, f0, df, impl=dd.FDD)      # default output mode
    plan.set_dm_list(g["dms"])
    nsamps = g["input"].shape[0]
    assert plan.out_length(nsamps) == g["fourier"].shape[1]
    out = plan.execute(g["input"], 8, 32)
    assert rel_l2(out, g["fourier"]) < REL_FP32


@pytest.mark.parametrize("nsamps,nch", [(40000, 64), (65536, 256), (2 ** 19, 1024), (1090000, 16),
                                         (16000, 16), (1160000, 16), (1200000, 16), (2350000, 16),
                                         (2410000, 16)])
def test_fdd_fourier_vs_oracle(dd, orc, nsamps, nch):
    """The whole-series Fourier output (the fork's default) against the oracle;
    the row transforms' m-point step (N = 16384 m) runs for odd m (3, 5, 67: the
    direct DFT, 67 prime), for even m <= 144 fused with the real-input pairing
    (one radix-2 split, mirror columns paired in LDS: 2, 34, 72 in the two
    workgroups per CU shape, 74, 144 in the one per CU shape) and for even
    m > 144 as two passes (148)."""
    rng = np.random.default_rng(nsamps)
    dt, f0, df = 54.6e-6, 1530.0, -0.3125 * 1024 / nch
    data = np.clip(np.rint(rng.normal(128, 20, size=(nsamps, nch)).astype(np.float32)), 0, 255).astype(np.uint8)
    delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    dms = np.array([0.0, 37.3, 511.0, 1023.5], f32) if nch == 1024 else np.linspace(0, 300, 9, dtype=f32)
    plan = dd.Plan(nch, dt, f0, df, impl=dd.FDD)
    plan.set_dm_list(dms)
    N = plan.out_length(nsamps)
    assert N == ((nsamps + plan.max_delay + 16383) // 16384) * 16384
    out = plan.execute(data, 8, 32)
    ref = orc.fdd_fourier(nch, delays, dms, nsamps, plan.max_delay, data)
    assert out.shape == ref.shape
    assert rel_l2(out, ref) < REL_FP32


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_fdd_fourier_matrix_forms_vs_oracle(dd, orc, prec):
    """The whole-series Fourier output with the bfloat16 and float16
    contraction forms (packed spectra, mirror bins, bin-block layout; m = 34,
    so the fused m-point step) against the oracle, at the forms' tolerances."""
    nsamps, nch = 2 ** 19, 1024
    rng = np.random.default_rng(nsamps + 7)
    dt, f0, df = 54.6e-6, 1530.0, -0.3125
    data = np.clip(np.rint(rng.normal(128, 20, size=(nsamps, nch)).astype(np.float32)), 0, 255).astype(np.uint8)
    delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    dms = np.array([0.0, 37.3, 511.0, 1023.5], f32)
    plan = dd.Plan(nch, dt, f0, df, impl=dd.FDD)
    plan.fdd_set_precision({"bf16": dd.FDD_BF16, "fp16": dd.FDD_FP16}[prec])
    plan.set_dm_list(dms)
    out = plan.execute(data, 8, 32)
    ref = orc.fdd_fourier(nch, delays, dms, nsamps, plan.max_delay, data)
    assert out.shape == ref.shape
    assert rel_l2(out, ref) < PREC_TOL[prec]


@pytest.mark.parametrize("nbits", [2, 4])
def test_fdd_subbyte_input_vs_oracle(dd, orc, nbits):
    """Sub-byte FDD input (an extension: the reference FDD reads 8 bits only):
    samples centred on (2^nbits - 1)/2, here on the many-beam geometry
    (512 ch 1500 MHz down in 0.78125 MHz, dt 81.92 us) with a shorter DM range."""
    from dedisp_tests_amd import synth
    nch, dt, f0, df = 512, 81.92e-6, 1500.0, -0.78125
    nsamps = 1 << 16
    delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    vmax = (1 << nbits) - 1
    s = synth.filterbank(nsamps, nch, nbits, vmax / 2.0, vmax / 4.0, seed=nbits,
                         pulses=[(150.0, 20000, 3, 2)], delays=delays)
    packed = dd.pack_samples(s, nbits)
    dms = np.linspace(0, 300, 13, dtype=f32)
    for prec in (dd.FDD_FP32, dd.FDD_BF16):
        plan = _fdd_plan(dd, nch, dt, f0, df, prec=prec)
        plan.set_dm_list(dms)
        out = plan.execute(packed, nbits, 32)
        ref = orc.fdd_timeseries(nch, delays, dms, nsamps, plan.max_delay, s.astype(np.uint8), nbits=nbits)
        assert rel_l2(out, ref) < REL_FP32


def test_fdd_host_pointers_many_dm_batches(dd):
    """Host-pointer FDD execute whose output spans several 256-DM batches of
    about 1 GB (downloads overlapped with later batches): identical to the
    device-pointer execute."""
    import torch
    nch = 64
    plan = dd.Plan(nch, 64e-6, 1581.8, -2.0, impl=dd.FDD)
    plan.fdd_set_output(dd.FDD_TIMESERIES)
    plan.fdd_set_precision(dd.FDD_BF16)
    plan.set_dm_list(np.linspace(0, 60, 600, dtype=np.float32))
    nsamps = 1 << 20
    md, ndm = plan.max_delay, plan.dm_count
    rng = np.random.default_rng(5)
    data = np.clip(np.rint(rng.normal(128, 20, size=(nsamps, nch))), 0, 255).astype(np.uint8)
    out = plan.execute(data, 8, 32)
    d_in = torch.from_numpy(data).cuda()
    d_out = torch.empty((ndm, nsamps - md), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    assert plan.execute_raw(nsamps, d_in, 8, d_out, 32, dd.DEVICE_POINTERS) == 0
    torch.cuda.synchronize()
    assert np.array_equal(out, d_out.cpu().numpy())


@pytest.mark.parametrize("prec", ["bf16", "fp16", "fp32-tile16"])
@pytest.mark.parametrize("nchunk_target", [1, 8, 9, 16, 23, 25, 33])
def test_fdd_tile_kernels_vs_oracle_chunk_counts(dd, orc, monkeypatch, prec, nchunk_target):
    """The 16 x 16 tile contraction (default for bfloat16 / float16) with
    mirror bins writing the inverse's packed spectrum (in the bin-block layout
    or in rows) or plain bins, without
    mirror bins, in each workgroup shape (the software-pipelined default, 8
    waves, 4 waves of 2 slots), and the 32 x 32 tile kernel, against the
    oracle for
    chunk counts that fill 1, 2 or 3 N-tiles and several chunk groups; 100
    channels pad to 128 rows."""
    p = {"bf16": dd.FDD_BF16, "fp16": dd.FDD_FP16, "fp32-tile16": dd.FDD_FP32}[prec]
    if prec == "fp32-tile16":
        monkeypatch.setenv("DEDISP_FDD_FP32_TILE16", "1")
    nch, dt, f0, df = 100, 64e-6, 1581.8, -1.0
    dms = np.linspace(0, 150, 70, dtype=f32)
    probe = _fdd_plan(dd, nch, dt, f0, df, nfft=8192)
    probe.set_dm_list(dms)
    good = (8192 - probe.max_delay) & ~1
    nsamps = probe.max_delay + nchunk_target * good - good // 3
    probe.close()
    rng = np.random.default_rng(nchunk_target)
    data = np.clip(np.rint(rng.normal(128, 20, size=(nsamps, nch))), 0, 255).astype(np.uint8)
    delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    ref = None
    outs = {}
    on = {"DEDISP_FDD_MIRROR": "1"}
    for mode, env in (("default", {}), ("mirror", on), ("mirror-rows", {**on, "DEDISP_FDD_BLOCKED": "0"}),
                      ("mirror-bins", {**on, "DEDISP_FDD_PACKED": "0"}),
                      ("direct", {"DEDISP_FDD_MIRROR": "0"}), ("qtable", {**on, "DEDISP_FDD_QTABLE": "1"}),
                      ("wg8", {**on, "DEDISP_FDD_M16CFG": "0"}), ("wg8-direct", {"DEDISP_FDD_M16CFG": "0",
                                                                                "DEDISP_FDD_MIRROR": "0"}),
                      ("wg4", {**on, "DEDISP_FDD_M16CFG": "1"}), ("wg4-direct", {"DEDISP_FDD_M16CFG": "1",
                                                                                "DEDISP_FDD_MIRROR": "0"}),
                      ("pipe", {**on, "DEDISP_FDD_M16CFG": "2"}),
                      ("pipe-direct", {"DEDISP_FDD_M16CFG": "2", "DEDISP_FDD_MIRROR": "0"}),
                      ("tile32", {"DEDISP_FDD_MFMA32": "1"})):
        for k in ("DEDISP_FDD_MIRROR", "DEDISP_FDD_MFMA32", "DEDISP_FDD_PACKED", "DEDISP_FDD_QTABLE",
                  "DEDISP_FDD_M16CFG", "DEDISP_FDD_BLOCKED"):
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        plan = _fdd_plan(dd, nch, dt, f0, df, nfft=8192, prec=p)
        plan.set_dm_list(dms)
        outs[mode] = plan.execute(data, 8, 32)
        if ref is None:
            ref = orc.fdd_timeseries(nch, delays, dms, nsamps, plan.max_delay, data)
        plan.close()
    tol = PREC_TOL["fp32" if prec == "fp32-tile16" else prec]
    errs = {mode: rel_l2(out, ref) for mode, out in outs.items()}
    assert all(e < tol for e in errs.values()), errs


# ---------------------------------------------------------------------------
# Maximum delays beyond 2^16 samples: split segments of 2^17 and 2^18
# ---------------------------------------------------------------------------

LONG_DELAY = dict(nch=128, dt=64e-6, f0=400.0, df=-1.0)   # 400 -> 273 MHz: ~465 samples per unit DM


def _long_delay_case(orc, nsamps, seed):
    g = LONG_DELAY
    delays = orc.delay_table(g["nch"], f32(g["dt"]), f32(g["f0"]), f32(g["df"]))
    dms = np.linspace(0.0, 172.0, 24, dtype=f32)
    rng = np.random.default_rng(seed)
    data = np.clip(np.rint(rng.normal(128, 20, size=(nsamps, g["nch"]))), 0, 255).astype(np.uint8)
    return delays, dms, data


@pytest.mark.parametrize("nfft,prec", [(0, "fp32"), (131072, "fp32"), (262144, "fp32"), (131072, "bf16"),
                                       (262144, "fp16")])
def test_fdd_long_delay_timeseries_vs_oracle(dd, orc, nfft, prec):
    """max_delay ~ 80000 samples: segments of 2^17 / 2^18 (split transforms),
    the automatic choice included, against the oracle."""
    g = LONG_DELAY
    nsamps = 230000
    delays, dms, data = _long_delay_case(orc, nsamps, 5)
    p = {"fp32": dd.FDD_FP32, "bf16": dd.FDD_BF16, "fp16": dd.FDD_FP16}[prec]
    plan = _fdd_plan(dd, g["nch"], g["dt"], g["f0"], g["df"], nfft=nfft, prec=p)
    if nfft == 0:
        plan.fdd_set_segmented(True, 0)
    plan.set_dm_list(dms)
    assert 78000 < plan.max_delay < 82000
    out = plan.execute(data, 8, 32)
    ref = orc.fdd_timeseries(g["nch"], delays, dms, nsamps, plan.max_delay, data)
    assert out.shape == ref.shape
    err = rel_l2(out, ref)
    assert err < PREC_TOL[prec], err
    plan.close()


def test_fdd_long_delay_fourier_vs_oracle(dd, orc):
    """The default Fourier-coefficient output with max_delay ~ 80000 (the
    circular shift-sum comes from split segments)."""
    g = LONG_DELAY
    nsamps = 150000
    delays, dms, data = _long_delay_case(orc, nsamps, 6)
    plan = dd.Plan(g["nch"], g["dt"], g["f0"], g["df"], impl=dd.FDD)
    plan.set_dm_list(dms)
    N = plan.out_length(nsamps)
    assert N == ((nsamps + plan.max_delay + 16383) // 16384) * 16384
    out = plan.execute(data, 8, 32)
    ref = orc.fdd_fourier(g["nch"], delays, dms, nsamps, plan.max_delay, data)
    assert rel_l2(out, ref) < REL_FP32
    plan.close()


@pytest.mark.parametrize("nfft", [131072, 262144])
def test_fdd_long_delay_batch_split_segments(dd, orc, nfft):
    """Batched execution with split segments (the chunks of both beams through
    one forward, contraction and inverse sequence): each beam equals a single
    execute at the same segment length, and the oracle."""
    g = LONG_DELAY
    plan = _fdd_plan(dd, g["nch"], g["dt"], g["f0"], g["df"], nfft=nfft, prec=dd.FDD_BF16)
    delays, dms, data = _long_delay_case(orc, 200000, 8)
    plan.set_dm_list(dms)
    beams = [np.ascontiguousarray(data[:120000]), np.ascontiguousarray(data[:200000])]
    outs = plan.execute_batch(beams, 8, 32)
    for b, o in zip(beams, outs):
        assert np.array_equal(o, plan.execute(b, 8, 32))
        ref = orc.fdd_timeseries(g["nch"], delays, dms, b.shape[0], plan.max_delay, b)
        assert rel_l2(o, ref) < PREC_TOL["bf16"]
    plan.close()
