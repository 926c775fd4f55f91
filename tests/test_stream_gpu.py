"""Streaming front end (dedisp_amd_stream_*) against one-shot execution.

After the n-th push of a beam the stream returns the dedispersed samples
t = n*chunk - max_delay + i of the beam's concatenated input. Every output with
t >= 0 reads only pushed samples, so it must equal the one-shot dedisp_execute
over the concatenation: bit-exact for TDD, within the FDD-FP32 tolerance for
FDD.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
f32 = np.float32


def _stitch(outs, chunk, hist):
    """Concatenated stream outputs for t >= 0 ([ndm, total - hist])."""
    full = np.concatenate(outs, axis=1)  # t = -hist .. n*chunk - hist - 1
    return full[:, hist:]


def test_tdd_stream_matches_one_shot(dd):
    from dedisp_tests_amd import synth
    nch, dt, f0, df = 256, 64e-6, 1581.8, -0.390625
    plan = dd.Plan(nch, dt, f0, df, impl=dd.TDD)
    plan.set_dm_list(np.arange(0, 200, 2, dtype=f32))
    hist = plan.max_delay
    chunk, npush = 500, 6  # chunk < max_delay (713) exercises a history longer than a chunk
    assert chunk < hist
    rng = np.random.default_rng(5)
    series = [rng.integers(0, 16, size=(chunk * npush, nch), dtype=np.uint32) for _ in range(2)]
    packed = [dd.pack_samples(s, 4) for s in series]
    st = dd.Stream(plan, nbeams=2, chunk_nsamps=chunk, in_nbits=4, out_nbits=32)
    assert st.history == hist
    outs = [[], []]
    for n in range(npush):  # interleave the beams: histories must stay apart
        for b in (1, 0):
            outs[b].append(st.push(b, packed[b][n * chunk:(n + 1) * chunk]))
    for b in range(2):
        got = _stitch(outs[b], chunk, hist)
        ref = plan.execute(packed[b], 4, 32)
        assert got.shape == ref.shape
        assert np.array_equal(got, ref)
    s = st.stats()
    assert s["pushes"] == 2 * npush and s["p99_ms"] >= s["p50_ms"] > 0
    # reset: the history is zero again, so the first push equals a fresh stream
    st.reset()
    again = st.push(0, packed[0][:chunk])
    assert np.array_equal(again, outs[0][0])
    st.close()


def test_tdd_stream_8bit_out_and_device_pointers(dd):
    import torch
    nch = 128
    plan = dd.Plan(nch, 64e-6, 1581.8, -0.78125, impl=dd.TDD)
    plan.set_dm_list(np.linspace(0, 150, 33, dtype=f32))
    hist, chunk, npush = plan.max_delay, 4096, 3
    rng = np.random.default_rng(9)
    data = rng.integers(0, 256, size=(chunk * npush, nch), dtype=np.uint8)
    st = dd.Stream(plan, nbeams=1, chunk_nsamps=chunk, in_nbits=8, out_nbits=8)
    d_out = torch.empty((plan.dm_count, chunk), dtype=torch.uint8, device="cuda")
    outs = []
    for n in range(npush):
        d_in = torch.from_numpy(np.ascontiguousarray(data[n * chunk:(n + 1) * chunk])).cuda()
        torch.cuda.synchronize()
        assert st.push_raw(0, d_in, d_out, dd.DEVICE_POINTERS) == 0
        outs.append(d_out.cpu().numpy().copy())
    got = _stitch(outs, chunk, hist)
    ref = plan.execute(data, 8, 8)
    assert np.array_equal(got, ref)


def test_fdd_stream_matches_oracle(dd, orc):
    nch, dt, f0, df = 64, 64e-6, 1581.8, -2.0
    plan = dd.Plan(nch, dt, f0, df, impl=dd.FDD)
    dms = np.linspace(0, 60, 9, dtype=f32)
    plan.set_dm_list(dms)
    hist, chunk, npush = plan.max_delay, 8000, 4
    rng = np.random.default_rng(3)
    data = np.clip(np.rint(rng.normal(128, 20, size=(chunk * npush, nch))), 0, 255).astype(np.uint8)
    st = dd.Stream(plan, nbeams=1, chunk_nsamps=chunk, in_nbits=8, out_nbits=32)
    outs = [st.push(0, data[n * chunk:(n + 1) * chunk]) for n in range(npush)]
    got = _stitch(outs, chunk, hist)
    delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    ref = orc.fdd_timeseries(nch, delays, dms, chunk * npush, hist, data)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err < 1e-4, err


def test_stream_admission_and_errors(dd):
    plan = dd.Plan(64, 64e-6, 1581.8, -2.0, impl=dd.TDD)
    import ctypes as C
    h = C.c_void_p(0)
    # no DM list yet
    assert dd.lib().dedisp_amd_stream_create(C.byref(h), plan.handle, 1, 1000, 8, 32) == dd.DEDISP_NO_DM_LIST_SET
    plan.set_dm_list(np.array([0.0, 10.0], f32))
    # far beyond any device's memory: refused by the admission check
    need = dd.lib().dedisp_amd_stream_bytes(plan.handle, 1 << 20, 1 << 20, 8, 32)
    assert need > (1 << 40)
    assert dd.lib().dedisp_amd_stream_create(C.byref(h), plan.handle, 1 << 20, 1 << 20, 8, 32) == \
        dd.DEDISP_MEM_ALLOC_FAILED
    assert dd.lib().dedisp_amd_stream_create(C.byref(h), plan.handle, 1, 1000, 3, 32) == \
        dd.DEDISP_UNSUPPORTED_IN_NBITS
    st = dd.Stream(plan, nbeams=2, chunk_nsamps=1000, in_nbits=8)
    buf = np.zeros((1000, 64), np.uint8)
    out = np.zeros((2, 1000), np.float32)
    assert st.push_raw(2, buf, out) == dd.DEDISP_INVALID_POINTER  # beam out of range
    plan.set_dm_list(np.array([0.0, 20.0], f32))                  # max_delay changes
    assert st.push_raw(0, buf, out) == dd.DEDISP_INVALID_PLAN


def test_stream_cfg5_geometry(dd, orc):
    """The streaming workload's geometry (1024 ch, 3000 trials 0..1500, 1 s
    chunks of 15625 samples at 64 us, max_delay 24849 > chunk) on two beams:
    three pushes each equal one-shot execution bit for bit, and two DM rows
    equal the oracle."""
    import torch
    from dedisp_tests_amd import synth
    g = synth.CFG5
    nch, dt, f0, df, chunk = g["nchans"], g["dt"], g["f0"], g["df"], g["chunk"]
    plan = dd.Plan(nch, dt, f0, df, impl=dd.TDD)
    dms = np.linspace(0.0, 1500.0, 3000, dtype=f32)
    plan.set_dm_list(dms)
    hist = plan.max_delay
    assert hist == 24849
    npush = 3
    st = dd.Stream(plan, nbeams=2, chunk_nsamps=chunk, in_nbits=8, out_nbits=32)
    d_out = torch.empty((3000, chunk), dtype=torch.float32, device="cuda")
    for beam in (0, 1):
        rng = np.random.default_rng(100 + beam)
        data = np.clip(np.rint(rng.normal(128, 20, size=(chunk * npush, nch))), 0, 255).astype(np.uint8)
        outs = []
        for n in range(npush):
            d_in = torch.from_numpy(np.ascontiguousarray(data[n * chunk:(n + 1) * chunk])).cuda()
            torch.cuda.synchronize()
            assert st.push_raw(beam, d_in, d_out, dd.DEVICE_POINTERS) == 0
            outs.append(d_out.cpu().numpy().copy())
        got = _stitch(outs, chunk, hist)
        d_all = torch.from_numpy(data).cuda()
        one = torch.empty((3000, chunk * npush - hist), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        assert plan.execute_raw(chunk * npush, d_all, 8, one, 32, dd.DEVICE_POINTERS) == 0
        assert np.array_equal(got, one.cpu().numpy())
        delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
        for i in (1, 2999):
            ref = orc.tdd(nch, delays, None, dms, chunk * npush, hist, data, 8, 32, first_dm=i, dm_count=1)
            assert np.array_equal(got[i], ref[0]), (beam, i)


@pytest.mark.parametrize("impl", ["TDD", "FDD_BF16"])
def test_stream_push_batch_matches_single_pushes(dd, impl):
    """Batched pushes (one chunk of several beams per call, through the plan's
    batched execution) give exactly the outputs of single pushes. The beam
    subsets differ between calls and repeat, so the TDD batch replays captured
    graphs over both of each beam's alternating history buffers."""
    import torch
    nch = 128
    if impl == "TDD":
        plan = dd.Plan(nch, 64e-6, 1581.8, -0.78125, impl=dd.TDD)
    else:
        plan = dd.Plan(nch, 64e-6, 1581.8, -0.78125, impl=dd.FDD)
        plan.fdd_set_precision(dd.FDD_BF16)
    plan.set_dm_list(np.linspace(0, 120, 45, dtype=f32))
    if impl != "TDD":
        plan.fdd_set_segmented(True, 8192)
    chunk, nb = 3000, 4
    ref_st = dd.Stream(plan, nbeams=nb, chunk_nsamps=chunk, in_nbits=8, out_nbits=32)
    bat_st = dd.Stream(plan, nbeams=nb, chunk_nsamps=chunk, in_nbits=8, out_nbits=32)
    rng = np.random.default_rng(17)
    ndm = plan.dm_count
    calls = [[0, 1, 2, 3], [2, 0], [1, 3, 0, 2], [0, 1, 2, 3], [3], [0, 1, 2, 3], [0, 1, 2, 3]]
    d_out = [torch.empty((ndm, chunk), dtype=torch.float32, device="cuda") for _ in range(nb)]
    one = torch.empty((ndm, chunk), dtype=torch.float32, device="cuda")
    for n, beams in enumerate(calls):
        chunks = [torch.from_numpy(np.clip(np.rint(rng.normal(128, 20, size=(chunk, nch))), 0, 255)
                                   .astype(np.uint8)).cuda() for _ in beams]
        torch.cuda.synchronize()
        assert bat_st.push_batch_raw(beams, chunks, [d_out[b] for b in beams], dd.DEVICE_POINTERS) == 0
        torch.cuda.synchronize()
        for b, c in zip(beams, chunks):
            assert ref_st.push_raw(b, c, one, dd.DEVICE_POINTERS) == 0
            torch.cuda.synchronize()
            assert torch.equal(d_out[b], one), (n, b)
    s = bat_st.stats()
    assert s["pushes"] == len(calls)
    assert s["samples_per_s"] > 0


def test_stream_push_batch_host_and_errors(dd):
    nch = 64
    plan = dd.Plan(nch, 64e-6, 1581.8, -2.0, impl=dd.TDD)
    plan.set_dm_list(np.linspace(0, 40, 11, dtype=f32))
    st = dd.Stream(plan, nbeams=3, chunk_nsamps=1000, in_nbits=8, out_nbits=16)
    ref = dd.Stream(plan, nbeams=3, chunk_nsamps=1000, in_nbits=8, out_nbits=16)
    rng = np.random.default_rng(2)
    for _ in range(3):
        chunks = [rng.integers(0, 256, size=(1000, nch), dtype=np.uint8) for _ in range(3)]
        outs = st.push_batch([2, 0, 1], [chunks[2], chunks[0], chunks[1]])
        for b, o in zip([2, 0, 1], outs):
            assert np.array_equal(o, ref.push(b, chunks[b])), b
    buf = np.zeros((1000, nch), np.uint8)
    out = np.zeros((11, 1000), np.uint16)
    assert st.push_batch_raw([1, 1], [buf, buf], [out, out]) == dd.DEDISP_INVALID_POINTER  # repeated beam
    assert st.push_batch_raw([3], [buf], [out]) == dd.DEDISP_INVALID_POINTER               # out of range
    assert st.push_batch_raw([], [], []) == 0


def test_stream_refuses_fdd_plan_switched_to_fourier(dd):
    """A stream switches its FDD plan to time series; switching the plan back to
    Fourier output afterwards (N floats per DM, longer than a chunk row) makes
    every push fail with DEDISP_INVALID_PLAN before anything runs, and the
    stream works again once time series is restored."""
    nch = 64
    plan = dd.Plan(nch, 64e-6, 1581.8, -2.0, impl=dd.FDD)
    plan.set_dm_list(np.linspace(0, 40, 5, dtype=f32))
    st = dd.Stream(plan, nbeams=2, chunk_nsamps=2000, in_nbits=8, out_nbits=32)
    buf = np.full((2000, nch), 128, np.uint8)
    out = np.zeros((5, 2000), np.float32)
    plan.fdd_set_output(dd.FDD_FOURIER)
    assert st.push_raw(0, buf, out) == dd.DEDISP_INVALID_PLAN
    assert st.push_batch_raw([0, 1], [buf, buf], [out, out.copy()]) == dd.DEDISP_INVALID_PLAN
    assert st.stats()["pushes"] == 0
    plan.fdd_set_output(dd.FDD_TIMESERIES)
    assert st.push_raw(0, buf, out) == 0
    st.close()
