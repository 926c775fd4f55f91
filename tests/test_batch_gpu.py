"""Batched beams (dedisp_amd_execute_batch) and the autotuner (dedisp_amd_autotune).

A batch must give every beam exactly what dedisp_execute on that beam alone
gives: bit-exact for TDD (one channel-sum launch over all beams), identical for
FDD (beams run one after another). With device pointers the TDD batch is a
captured HIP graph; it is replayed several times with other device work and
synchronisations in between, with the input contents changed in place, and
every replay is checked against single executes.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
f32 = np.float32


def _packed(dd, rng, nsamps, nch, nbits):
    s = rng.integers(0, 1 << nbits, size=(nsamps, nch), dtype=np.uint32)
    return dd.pack_samples(s, nbits)


def test_tdd_batch_host_pointers_match_single(dd):
    nch, nbits = 256, 2
    plan = dd.Plan(nch, 64e-6, 1581.8, -0.390625, impl=dd.TDD)
    plan.set_dm_list(np.arange(0, 120, 1.1, dtype=f32))
    md = plan.max_delay
    rng = np.random.default_rng(31)
    # uneven lengths, one with no output samples (nsamps == max_delay)
    lengths = [3000, 5137, md, md + 1, 9001]
    beams = [_packed(dd, rng, n, nch, nbits) for n in lengths]
    outs = plan.execute_batch(beams, nbits, 32)
    for i, (b, n) in enumerate(zip(beams, lengths)):
        assert outs[i].shape == (plan.dm_count, n - md)
        if n > md:
            assert np.array_equal(outs[i], plan.execute(b, nbits, 32)), i


@pytest.mark.parametrize("nbits,out_nbits", [(4, 8), (8, 32), (1, 16)])
def test_tdd_batch_widths(dd, nbits, out_nbits):
    nch = 128
    plan = dd.Plan(nch, 64e-6, 1500.0, -0.78125, impl=dd.TDD)
    plan.set_dm_list(np.linspace(0, 90, 37, dtype=f32))
    rng = np.random.default_rng(nbits)
    lengths = [4000, 2500, 7777]
    beams = [_packed(dd, rng, n, nch, nbits) for n in lengths]
    outs = plan.execute_batch(beams, nbits, out_nbits)
    for i, b in enumerate(beams):
        assert np.array_equal(outs[i], plan.execute(b, nbits, out_nbits)), i


def test_tdd_batch_graph_replays(dd):
    import torch
    nch, nbits = 256, 4
    plan = dd.Plan(nch, 64e-6, 1581.8, -0.390625, impl=dd.TDD)
    plan.set_dm_list(np.arange(0, 80, 0.7, dtype=f32))
    md, ndm = plan.max_delay, plan.dm_count
    rng = np.random.default_rng(5)
    lengths = [6000, 3100, 12001, 4444]
    d_in = [torch.zeros((n, nch * nbits // 8), dtype=torch.uint8, device="cuda") for n in lengths]
    d_out = [torch.empty((ndm, n - md), dtype=torch.float32, device="cuda") for n in lengths]
    one = [torch.empty((ndm, n - md), dtype=torch.float32, device="cuda") for n in lengths]
    for rep in range(4):
        # new contents in the same buffers: a replay must read them
        for t, n in zip(d_in, lengths):
            t.copy_(torch.from_numpy(_packed(dd, rng, n, nch, nbits)))
        torch.cuda.synchronize()
        assert plan.execute_batch_raw(lengths, d_in, nbits, d_out, 32, dd.DEVICE_POINTERS) == 0
        torch.cuda.synchronize()
        # other device work between replays
        x = torch.randn(1 << 20, device="cuda")
        (x * 2).sum().item()
        for t, o, n in zip(d_in, one, lengths):
            assert plan.execute_raw(n, t, nbits, o, 32, dd.DEVICE_POINTERS) == 0
        torch.cuda.synchronize()
        for i in range(len(lengths)):
            assert torch.equal(d_out[i], one[i]), (rep, i)
    # a different batch (one length changed) is captured anew
    lengths2 = [6000, 3100, 9000, 4444]
    d_in2 = [d_in[0], d_in[1], torch.from_numpy(_packed(dd, rng, 9000, nch, nbits)).cuda(), d_in[3]]
    d_out2 = [d_out[0], d_out[1], torch.empty((ndm, 9000 - md), dtype=torch.float32, device="cuda"), d_out[3]]
    torch.cuda.synchronize()
    assert plan.execute_batch_raw(lengths2, d_in2, nbits, d_out2, 32, dd.DEVICE_POINTERS) == 0
    ref = torch.empty_like(d_out2[2])
    assert plan.execute_raw(9000, d_in2[2], nbits, ref, 32, dd.DEVICE_POINTERS) == 0
    torch.cuda.synchronize()
    assert torch.equal(d_out2[2], ref)
    assert torch.equal(d_out2[0], one[0])


def test_tdd_batch_weighted_killmask_runs_beam_by_beam(dd):
    nch, nbits = 64, 8
    plan = dd.Plan(nch, 64e-6, 1581.8, -2.0, impl=dd.TDD)
    plan.set_dm_list(np.arange(0, 40, 2.0, dtype=f32))
    km = np.ones(nch, np.int32)
    km[3] = 2
    plan.set_killmask(km)
    rng = np.random.default_rng(2)
    beams = [_packed(dd, rng, n, nch, nbits) for n in (2000, 3500)]
    outs = plan.execute_batch(beams, nbits, 32)
    for i, b in enumerate(beams):
        assert np.array_equal(outs[i], plan.execute(b, nbits, 32))


def test_batch_errors(dd):
    plan = dd.Plan(64, 64e-6, 1581.8, -2.0, impl=dd.TDD)
    beams = [np.zeros((1000, 64), np.uint8)]
    out = [np.zeros((1, 1000), np.float32)]
    assert plan.execute_batch_raw([1000], beams, 8, out, 32) == dd.DEDISP_NO_DM_LIST_SET
    plan.set_dm_list(np.array([0.0, 30.0], f32))
    assert plan.execute_batch_raw([10], beams, 8, out, 32) == dd.DEDISP_TOO_FEW_NSAMPS
    assert plan.execute_batch_raw([1000], beams, 3, out, 32) == dd.DEDISP_UNSUPPORTED_IN_NBITS
    assert plan.execute_batch_raw([1000], [None], 8, out, 32) == dd.DEDISP_INVALID_POINTER
    assert plan.execute_batch_raw([], [], 8, [], 32) == 0


def _u8_beams(rng, lengths, nch):
    return [np.clip(np.rint(rng.normal(128, 20, size=(n, nch))), 0, 255).astype(np.uint8) for n in lengths]


def test_fdd_batch_matches_single(dd):
    """Chunks of all beams in one forward / contraction / inverse sequence:
    every beam equals its single execute at the same segment length."""
    nch = 64
    plan = dd.Plan(nch, 64e-6, 1581.8, -2.0, impl=dd.FDD)
    plan.fdd_set_output(dd.FDD_TIMESERIES)
    plan.set_dm_list(np.linspace(0, 60, 9, dtype=f32))
    plan.fdd_set_segmented(True, 8192)
    rng = np.random.default_rng(4)
    md = plan.max_delay
    # uneven lengths; one beam without outputs (nsamps == max_delay)
    beams = _u8_beams(rng, (9000, 14000, md, 6000), nch)
    outs = plan.execute_batch(beams, 8, 32)
    for i, b in enumerate(beams):
        assert outs[i].shape == (plan.dm_count, b.shape[0] - md)
        if b.shape[0] > md:
            assert np.array_equal(outs[i], plan.execute(b, 8, 32)), i


@pytest.mark.parametrize("prec", ["FDD_FP32", "FDD_BF16", "FDD_FP16"])
def test_fdd_batch_groups_device_pointers(dd, prec, monkeypatch):
    """Several groups of beams (the chunks per group capped at 5, so groups
    split between beams and a beam longer than the cap forms its own group),
    100 channels (padded to 112 in the contraction), every contraction form,
    device pointers: each beam equals its single execute bit for bit."""
    import torch
    nch = 100
    plan = dd.Plan(nch, 64e-6, 1581.8, -1.5, impl=dd.FDD)
    plan.fdd_set_output(dd.FDD_TIMESERIES)
    plan.fdd_set_precision(getattr(dd, prec))
    plan.set_dm_list(np.linspace(0, 90, 70, dtype=f32))
    plan.fdd_set_segmented(True, 8192)
    md, ndm = plan.max_delay, plan.dm_count
    rng = np.random.default_rng(12)
    lengths = [20000, 9000, 61000, 15000, 30000, 12001]
    host = _u8_beams(rng, lengths, nch)
    d_in = [torch.from_numpy(h).cuda() for h in host]
    d_out = [torch.empty((ndm, n - md), dtype=torch.float32, device="cuda") for n in lengths]
    monkeypatch.setenv("DEDISP_FDD_GROUP_CHUNKS", "5")
    torch.cuda.synchronize()
    assert plan.execute_batch_raw(lengths, d_in, 8, d_out, 32, dd.DEVICE_POINTERS) == 0
    torch.cuda.synchronize()
    monkeypatch.delenv("DEDISP_FDD_GROUP_CHUNKS")
    for i, n in enumerate(lengths):
        one = torch.empty_like(d_out[i])
        assert plan.execute_raw(n, d_in[i], 8, one, 32, dd.DEVICE_POINTERS) == 0
        torch.cuda.synchronize()
        assert torch.equal(d_out[i], one), i


def test_fdd_batch_subbyte_auto_segment(dd, orc):
    """4-bit input, host pointers, segment length chosen for the batch: each
    beam within the FDD-FP32 tolerance of the oracle."""
    nch, dt, f0, df = 64, 64e-6, 1581.8, -2.0
    plan = dd.Plan(nch, dt, f0, df, impl=dd.FDD)
    plan.fdd_set_output(dd.FDD_TIMESERIES)
    dms = np.linspace(0, 60, 9, dtype=f32)
    plan.set_dm_list(dms)
    rng = np.random.default_rng(21)
    lengths = [12000, 30001, 7000]
    raw = [rng.integers(0, 16, size=(n, nch), dtype=np.uint32) for n in lengths]
    beams = [dd.pack_samples(r, 4) for r in raw]
    outs = plan.execute_batch(beams, 4, 32)
    delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    for i, (r, n) in enumerate(zip(raw, lengths)):
        ref = orc.fdd_timeseries(nch, delays, dms, n, plan.max_delay, r.astype(np.uint8), nbits=4)
        err = np.linalg.norm(outs[i] - ref) / np.linalg.norm(ref)
        assert err < 1e-4, (i, err)


def test_fdd_batch_fourier_output_runs_beam_by_beam(dd):
    nch = 64
    plan = dd.Plan(nch, 64e-6, 1581.8, -2.0, impl=dd.FDD)
    plan.set_dm_list(np.linspace(0, 30, 5, dtype=f32))
    rng = np.random.default_rng(9)
    beams = _u8_beams(rng, (5000, 7000), nch)
    outs = plan.execute_batch(beams, 8, 32)
    for i, b in enumerate(beams):
        assert np.array_equal(outs[i], plan.execute(b, 8, 32)), i


def test_autotune_tdd_keeps_results_exact(dd, orc):
    nch, nbits = 256, 4
    dt, f0, df = 64e-6, 1581.8, -0.390625
    plan = dd.Plan(nch, dt, f0, df, impl=dd.TDD)
    dms = np.arange(0, 150, 1.3, dtype=f32)
    plan.set_dm_list(dms)
    r = plan.autotune(1 << 16, nbits, 32)
    assert r["candidates"] >= 5 and r["best_ms"] > 0
    assert r["samples_per_lane"] % 8 == 0 and r["dms_per_wave"] in (1, 2, 4) and r["waves"] in (4, 8)
    rng = np.random.default_rng(8)
    data = _packed(dd, rng, 20000, nch, nbits)
    out = plan.execute(data, nbits, 32)
    delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    ref = orc.tdd(nch, delays, None, dms, 20000, plan.max_delay, data, nbits, 32)
    assert np.array_equal(out, ref)


def test_autotune_fdd_segment_length(dd, orc):
    nch, dt, f0, df = 64, 64e-6, 1581.8, -2.0
    plan = dd.Plan(nch, dt, f0, df, impl=dd.FDD)
    plan.fdd_set_output(dd.FDD_TIMESERIES)
    dms = np.linspace(0, 60, 9, dtype=f32)
    plan.set_dm_list(dms)
    r = plan.autotune(40000, 8, 32)
    assert r["nfft"] in (8192, 16384, 32768, 65536) and r["nfft"] > plan.max_delay
    assert r["candidates"] >= 1 and r["best_ms"] > 0
    rng = np.random.default_rng(6)
    data = np.clip(np.rint(rng.normal(128, 20, size=(30000, nch))), 0, 255).astype(np.uint8)
    out = plan.execute(data, 8, 32)
    delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    ref = orc.fdd_timeseries(nch, delays, dms, 30000, plan.max_delay, data)
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1e-4


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16"])
def test_autotune_fdd_contraction_form(dd, orc, monkeypatch, prec):
    """The autotuner also times the contraction forms of the plan's precision
    (workgroup shape, mirror bins, spectra layout) at the chosen segment
    length; the kept form is one of them and the executes after it stay
    within the precision's tolerance of the oracle."""
    for k in ("DEDISP_FDD_MIRROR", "DEDISP_FDD_MFMA32", "DEDISP_FDD_PACKED", "DEDISP_FDD_QTABLE",
              "DEDISP_FDD_M16CFG", "DEDISP_FDD_BLOCKED", "DEDISP_FDD_FP32_TILE16", "DEDISP_FDD_VALU"):
        monkeypatch.delenv(k, raising=False)
    nch, dt, f0, df = 96, 64e-6, 1581.8, -2.0
    plan = dd.Plan(nch, dt, f0, df, impl=dd.FDD)
    plan.fdd_set_output(dd.FDD_TIMESERIES)
    plan.fdd_set_precision({"fp32": dd.FDD_FP32, "bf16": dd.FDD_BF16, "fp16": dd.FDD_FP16}[prec])
    dms = np.linspace(0, 60, 40, dtype=f32)
    plan.set_dm_list(dms)
    nsamps = 120000
    r = plan.autotune(nsamps, 8, 32)
    forms = {(3, -1, -1), (0, 0, -1), (0, 1, -1)} if prec == "fp32" else \
        {(2, 1, 1), (2, 1, 0), (2, 0, -1), (0, 1, -1), (0, 0, -1), (3, -1, -1)}
    assert (r["fdd_form"], r["fdd_mirror"], r["fdd_layout"]) in forms, r
    assert r["candidates"] >= len(forms) + 1 and r["best_ms"] > 0
    rng = np.random.default_rng(9)
    data = np.clip(np.rint(rng.normal(128, 20, size=(nsamps, nch))), 0, 255).astype(np.uint8)
    out = plan.execute(data, 8, 32)
    delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    ref = orc.fdd_timeseries(nch, delays, dms, nsamps, plan.max_delay, data)
    tol = 1e-3 if prec == "fp16" else 1e-4
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < tol
    plan.close()


def test_batch_uneven_beams_512ch_2bit(dd, orc):
    """The batched-beams geometry (512 ch 1500 -> 1100 MHz, 81.92 us, 2-bit,
    1024 DMs 0..2000, max_delay 38585) on 6 beams of log-uniform length in
    [2^17, 2^18], 8-bit output, device pointers: every beam equals its single
    execute bit for bit, and two DM rows of one beam equal the oracle."""
    import torch
    from dedisp_tests_amd import synth
    g = synth.CFG4
    nch, nbits = g["nchans"], 2
    plan = dd.Plan(nch, g["dt"], g["f0"], g["df"], impl=dd.TDD)
    dms = np.linspace(0, 2000, 1024, dtype=f32)
    plan.set_dm_list(dms)
    md, ndm = plan.max_delay, plan.dm_count
    assert md == 38585
    rng = np.random.default_rng(44)
    lengths = [int(2 ** rng.uniform(17, 18)) for _ in range(6)]
    host = [_packed(dd, rng, n, nch, nbits) for n in lengths]
    d_in = [torch.from_numpy(h).cuda() for h in host]
    d_out = [torch.empty((ndm, n - md), dtype=torch.uint8, device="cuda") for n in lengths]
    torch.cuda.synchronize()
    assert plan.execute_batch_raw(lengths, d_in, nbits, d_out, 8, dd.DEVICE_POINTERS) == 0
    torch.cuda.synchronize()
    for i, n in enumerate(lengths):
        one = torch.empty_like(d_out[i])
        assert plan.execute_raw(n, d_in[i], nbits, one, 8, dd.DEVICE_POINTERS) == 0
        torch.cuda.synchronize()
        assert torch.equal(d_out[i], one), i
    delays = orc.delay_table(nch, f32(g["dt"]), f32(g["f0"]), f32(g["df"]))
    got = d_out[2].cpu().numpy()
    for r in (3, 1000):
        ref = orc.tdd(nch, delays, None, dms, lengths[2], md, host[2], nbits, 8, first_dm=r, dm_count=1)
        assert np.array_equal(got[r], ref[0]), r


def test_fdd_batch_graph_replays(dd):
    """Device-pointer FDD batches are captured once and replayed: new input
    contents each time, single executes (which reuse the plan's work buffers)
    and unrelated device work in between, 100 channels so the graph also
    clears the padded spectrum rows; every replay equals single executes."""
    import torch
    nch = 100
    plan = dd.Plan(nch, 64e-6, 1581.8, -1.5, impl=dd.FDD)
    plan.fdd_set_output(dd.FDD_TIMESERIES)
    plan.fdd_set_precision(dd.FDD_BF16)
    plan.set_dm_list(np.linspace(0, 90, 40, dtype=f32))
    plan.fdd_set_segmented(True, 8192)
    md, ndm = plan.max_delay, plan.dm_count
    rng = np.random.default_rng(23)
    lengths = [9000, 21000, 5000]
    d_in = [torch.zeros((n, nch), dtype=torch.uint8, device="cuda") for n in lengths]
    d_out = [torch.empty((ndm, n - md), dtype=torch.float32, device="cuda") for n in lengths]
    one = [torch.empty_like(o) for o in d_out]
    for rep in range(4):
        for t, n in zip(d_in, lengths):
            t.copy_(torch.from_numpy(_u8_beams(rng, (n,), nch)[0]))
        torch.cuda.synchronize()
        assert plan.execute_batch_raw(lengths, d_in, 8, d_out, 32, dd.DEVICE_POINTERS) == 0
        torch.cuda.synchronize()
        x = torch.randn(1 << 20, device="cuda")
        (x * 2).sum().item()
        for t, o, n in zip(d_in, one, lengths):
            assert plan.execute_raw(n, t, 8, o, 32, dd.DEVICE_POINTERS) == 0
        torch.cuda.synchronize()
        for i in range(len(lengths)):
            assert torch.equal(d_out[i], one[i]), (rep, i)
