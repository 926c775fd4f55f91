"""TDD parity on the GPU: the HIP path through the C ABI against the golden
fixtures and the oracle, bit-exact (integer sums, float epilogue)."""
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).resolve().parent / "golden"
f32 = np.float32
pytestmark = pytest.mark.gpu


def _plan(dd, nch, dt, f0, df, impl=None):
    return dd.Plan(nch, dt, f0, df, impl=dd.TDD if impl is None else impl)


def _cases():
    g = np.load(GOLD / "tdd_small.npz")
    return [str(c) for c in g["cases"]]


@pytest.mark.parametrize("case", _cases())
def test_tdd_golden_fixture(dd, case):
    g = np.load(GOLD / "tdd_small.npz")
    nch, dt, f0, df = g["geometry"]
    in_nbits = int(case.split("_")[0][1:])
    out_nbits = int(case.split("_")[1][1:])
    plan = _plan(dd, int(nch), dt, f0, df)
    plan.set_dm_list(g[case + "_dms"])
    plan.set_killmask(g[case + "_km"])
    out = plan.execute(g[case + "_in"], in_nbits, out_nbits)
    assert np.array_equal(out, g[case + "_out"])


def _random_case(dd, orc, nbits, nch, nsamps, dms, seed, killmask=None, out_nbits=32,
                 geom=(64e-6, 1581.8, -0.390625)):
    rng = np.random.default_rng(seed)
    hi = (1 << nbits) if nbits < 32 else (1 << 32)
    s = rng.integers(0, hi, size=(nsamps, nch), dtype=np.uint64).astype(np.uint32)
    data = dd.pack_samples(s, nbits)
    plan = _plan(dd, nch, *geom)
    plan.set_dm_list(dms)
    if killmask is not None:
        plan.set_killmask(killmask)
    out = plan.execute(data, nbits, out_nbits)
    delays = orc.delay_table(nch, f32(geom[0]), f32(geom[1]), f32(geom[2]))
    ref = orc.tdd(nch, delays, killmask, dms, nsamps, plan.max_delay, data, nbits, out_nbits)
    return out, ref


@pytest.mark.parametrize("nbits", [1, 2, 4, 8, 16, 32])
@pytest.mark.parametrize("out_nbits", [8, 32])
def test_tdd_random_vs_oracle(dd, orc, nbits, out_nbits):
    dms = np.arange(0, 70, 1.7, dtype=f32)
    out, ref = _random_case(dd, orc, nbits, 256, 5000 + 37, dms, seed=nbits * 7 + out_nbits,
                            out_nbits=out_nbits)
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("nbits", [2, 4, 8])
def test_tdd_killmask_binary_and_weighted(dd, orc, nbits):
    nch = 128
    dms = np.linspace(0, 40, 23, dtype=f32)
    km = np.ones(nch, np.int32)
    km[::5] = 0
    out, ref = _random_case(dd, orc, nbits, nch, 3000, dms, seed=11, killmask=km)
    assert np.array_equal(out, ref)
    km2 = km.copy()
    km2[3] = 2
    km2[7] = -1   # int * unsigned wraps mod 2^32 in the reference
    out, ref = _random_case(dd, orc, nbits, nch, 3000, dms, seed=12, killmask=km2)
    assert np.array_equal(out, ref)
    out, ref = _random_case(dd, orc, nbits, nch, 3000, dms, seed=13, killmask=np.zeros(nch, np.int32))
    assert np.array_equal(out, ref)
    assert np.all(out == 0)


@pytest.mark.parametrize("tiling", ["16,2,4,64", "16,4,4,64", "8,4,4,32", "16,2,8,64"])
def test_tdd_4bit_tilings(dd, orc, monkeypatch, tiling):
    monkeypatch.setenv("DEDISP_TDD_TILING", tiling)
    dms = np.arange(0, 90, 0.9, dtype=f32)
    out, ref = _random_case(dd, orc, 4, 512, 9000, dms, seed=21, geom=(54.6e-6, 1530.0, -0.625))
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("shape", ["4,4", "2,8"])
@pytest.mark.parametrize("cap", [64, 3, 1])
def test_tdd_runs_kernel_vs_oracle(dd, orc, monkeypatch, cap, shape):
    # The runs-layout kernel (4-bit, samples_per_lane 28; 4 DMs x 4 waves or
    # 2 DMs x 8 waves): 100 DMs (the last block of 16 partly filled), 512
    # channels; cap = channels per staged chunk (64: clamped to 32; 3: odd
    # chunks, byte lanes flushed across chunk boundaries; 1: one channel per
    # chunk).
    monkeypatch.setenv("DEDISP_TDD_TILING", f"28,{shape},{cap}")
    dms = np.arange(0, 90, 0.9, dtype=f32)
    out, ref = _random_case(dd, orc, 4, 512, 9000, dms, seed=31 + cap, geom=(54.6e-6, 1530.0, -0.625))
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("shape", ["4,4", "2,8"])
def test_tdd_runs_kernel_wide_windows(dd, orc, monkeypatch, shape):
    # DM steps of 2.9 at DM ~1000 over 1024 channels 1530 MHz down by 0.3125:
    # a block's windows span up to ~900 samples, so chunks hold few channels
    # and the positions per chunk vary along the band.
    monkeypatch.setenv("DEDISP_TDD_TILING", f"28,{shape},32")
    dms = np.arange(950.0, 950.0 + 40 * 2.9, 2.9, dtype=f32)
    out, ref = _random_case(dd, orc, 4, 1024, 26000, dms, seed=41, geom=(54.6e-6, 1530.0, -0.3125))
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("nch,out_nbits", [(38, 8), (130, 16), (202, 32)])
def test_tdd_runs_kernel_ragged_and_killmask(dd, orc, monkeypatch, nch, out_nbits):
    monkeypatch.setenv("DEDISP_TDD_TILING", "28,4,4,64")
    km = np.ones(nch, np.int32)
    km[::7] = 0
    dms = np.linspace(0, 45, 19, dtype=f32)
    out, ref = _random_case(dd, orc, 4, nch, 4000 + nch, dms, seed=50 + nch, killmask=km, out_nbits=out_nbits)
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("tiling", ["8,2,4,64", "16,2,4,64", "8,4,4,64", "8,2,8,32"])
def test_tdd_8bit_tilings(dd, orc, monkeypatch, tiling):
    monkeypatch.setenv("DEDISP_TDD_TILING", tiling)
    dms = np.arange(0, 50, 0.5, dtype=f32)
    # 300 channels of 255: 16-bit lanes are widened every staged chunk
    out, ref = _random_case(dd, orc, 8, 300, 7000, dms, seed=22, geom=(64e-6, 1581.8, -0.390625))
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("nbits,tiling", [(2, "16,4,4,32"), (2, "16,2,8,32"), (2, "32,2,4,16"),
                                          (1, "32,4,4,32"), (1, "32,2,8,32")])
def test_tdd_subbyte_tilings(dd, orc, monkeypatch, nbits, tiling):
    # 102 DMs (the last block is partly filled: padded groups repeat the last DM) and 200
    # channels of 2-bit / 1-bit samples (byte lanes flushed every 84 / 254)
    monkeypatch.setenv("DEDISP_TDD_TILING", tiling)
    dms = np.arange(0, 61, 0.6, dtype=f32)
    out, ref = _random_case(dd, orc, nbits, 200, 6000, dms, seed=23 + nbits, geom=(64e-6, 1500.0, -0.78125))
    assert np.array_equal(out, ref)


def test_tdd_cfg1_full_bit_exact(dd, orc):
    # 256 channels 1581.8 MHz down by 0.390625 MHz, dt 64 us, 65536 8-bit samples,
    # one DM-100 pulse, 256 DMs 0..255: bit-exact float output.
    from dedisp_tests_amd import synth
    nch, dt, f0, df = 256, 64e-6, 1581.8, -0.390625
    delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    s = synth.filterbank(65536, nch, 8, 128.0, 20.0, seed=1, pulses=[(100.0, 30000, 4, 30)], delays=delays)
    data = dd.pack_samples(s, 8)
    dms = np.arange(256, dtype=f32)
    plan = _plan(dd, nch, dt, f0, df)
    plan.set_dm_list(dms)
    assert plan.max_delay == 918
    out = plan.execute(data, 8, 32)
    ref = orc.tdd(nch, delays, None, dms, 65536, 918, data, 8, 32)
    assert np.array_equal(out, ref)
    # the pulse peaks at DM 100
    assert int(np.argmax(out.max(axis=1))) == 100


def test_tdd_adv_strides_and_guru_subset(dd, orc):
    rng = np.random.default_rng(5)
    nch, nsamps, nbits = 64, 2500, 4
    s = rng.integers(0, 16, size=(nsamps, nch)).astype(np.uint32)
    packed = dd.pack_samples(s, nbits)                  # 32 bytes per row
    padded = np.zeros((nsamps, 48), np.uint8)           # in_stride 48
    padded[:, :32] = packed
    dms = np.linspace(0, 30, 17, dtype=f32)
    plan = _plan(dd, nch, 64e-6, 1581.8, -2.0)
    plan.set_dm_list(dms)
    nsc = nsamps - plan.max_delay
    out_stride = (nsc + 13) * 4
    out = np.full((17, out_stride // 4), -1.0, f32)
    rc = plan.execute_adv(nsamps, padded, nbits, 48, out, 32, out_stride)
    assert rc == 0
    delays = orc.delay_table(nch, f32(64e-6), f32(1581.8), f32(-2.0))
    ref = orc.tdd(nch, delays, None, dms, nsamps, plan.max_delay, packed, nbits, 32)
    assert np.array_equal(out[:, :nsc], ref)
    assert np.all(out[:, nsc:] == -1.0)                 # padding untouched
    # guru: DMs 5..11 only
    out2 = np.zeros((7, nsc), f32)
    rc = plan.execute_guru(nsamps, packed, nbits, 32, out2, 32, nsc * 4, 5, 7)
    assert rc == 0
    assert np.array_equal(out2, ref[5:12])


@pytest.mark.parametrize("nbits", [2, 4, 8])
@pytest.mark.parametrize("nch,pad", [(12, 3), (100, 1), (300, 0), (300, 2)])
def test_tdd_ragged_channels_and_unaligned_strides(dd, orc, nbits, nch, pad):
    # channel counts (filling whole bytes, as the input layout requires) that
    # leave part of a sample row's last 32-bit word unused, and row strides
    # that are not a multiple of 4 bytes (the transpose's byte loads); the
    # padding bytes must be ignored
    rng = np.random.default_rng(nch * 10 + nbits + pad)
    nsamps = 3000 + 77
    s = rng.integers(0, 1 << nbits, size=(nsamps, nch)).astype(np.uint32)
    packed = dd.pack_samples(s, nbits)
    rb = packed.shape[1]
    padded = np.full((nsamps, rb + pad), 0xA5, np.uint8)
    padded[:, :rb] = packed
    dms = np.linspace(0, 40, 23, dtype=f32)
    plan = _plan(dd, nch, 64e-6, 1581.8, -0.390625)
    plan.set_dm_list(dms)
    nsc = nsamps - plan.max_delay
    out = np.zeros((len(dms), nsc), f32)
    assert plan.execute_adv(nsamps, padded, nbits, rb + pad, out, 32, nsc * 4) == 0
    delays = orc.delay_table(nch, f32(64e-6), f32(1581.8), f32(-0.390625))
    ref = orc.tdd(nch, delays, None, dms, nsamps, plan.max_delay, packed, nbits, 32)
    assert np.array_equal(out, ref)


def test_tdd_validation_errors(dd):
    plan = _plan(dd, 64, 64e-6, 1581.8, -2.0)
    buf = np.zeros((100, 64), np.uint8)
    out = np.zeros(100000, np.float32)
    assert plan.execute_raw(100, buf, 8, out, 32) == dd.DEDISP_NO_DM_LIST_SET
    plan.set_dm_list(np.array([0.0, 50.0], f32))
    md = plan.max_delay
    assert md > 10
    assert plan.execute_raw(100, None, 8, out, 32) == dd.DEDISP_INVALID_POINTER
    # nsamps < max_delay: execute derives out_stride with the same unsigned wrap,
    # so the sample check reports it; an explicit stride fails the stride check
    # first (reference order, src/tdd/TDDPlan.cpp:171-180)
    assert plan.execute_raw(md - 1, buf, 8, out, 32) == dd.DEDISP_TOO_FEW_NSAMPS
    assert plan.execute_adv(md - 1, buf, 8, 64, out, 32, 400) == dd.DEDISP_INVALID_STRIDE
    assert plan.execute_raw(md + 10, buf, 3, out, 32) == dd.DEDISP_UNSUPPORTED_IN_NBITS
    assert plan.execute_raw(md + 10, buf, 8, out, 12) == dd.DEDISP_UNSUPPORTED_OUT_NBITS
    assert plan.execute_adv(md + 10, buf, 8, 10, out, 32, 40) == dd.DEDISP_INVALID_STRIDE
    assert plan.execute_raw(md + 10, buf, 8, out, 32, dd.ASYNC | dd.WAIT) == dd.DEDISP_INVALID_FLAG_COMBINATION
    assert plan.execute_guru(md + 10, buf, 8, 64, out, 32, 40, 1, 5) != dd.DEDISP_NO_ERROR
    # nchans over the reference's limit
    with pytest.raises(dd.DedispError) as e:
        dd.Plan(8193, 64e-6, 1581.8, -0.1, impl=dd.TDD)
    assert e.value.code == dd.DEDISP_NCHANS_EXCEEDS_LIMIT


def test_tdd_gulp_size_api(dd):
    plan = _plan(dd, 64, 64e-6, 1581.8, -2.0)
    assert dd.lib().dedisp_get_gulp_size(plan.handle) == 65536
    assert dd.lib().dedisp_set_gulp_size(plan.handle, 1000) == 0
    assert dd.lib().dedisp_get_gulp_size(plan.handle) == 1000


def test_tdd_legacy_selector_runs_tdd(dd, orc):
    dms = np.linspace(0, 20, 9, dtype=f32)
    rng = np.random.default_rng(9)
    s = rng.integers(0, 256, size=(1500, 64)).astype(np.uint32)
    data = dd.pack_samples(s, 8)
    plan = dd.Plan(64, 64e-6, 1581.8, -2.0, impl=dd.DEDISP)
    plan.set_dm_list(dms)
    out = plan.execute(data, 8, 16)
    delays = orc.delay_table(64, f32(64e-6), f32(1581.8), f32(-2.0))
    assert np.array_equal(out, orc.tdd(64, delays, None, dms, 1500, plan.max_delay, data, 8, 16))


def test_tdd_device_pointers(dd, orc):
    import torch
    rng = np.random.default_rng(31)
    nch, nsamps = 256, 20000
    s = rng.integers(0, 16, size=(nsamps, nch)).astype(np.uint32)
    data = dd.pack_samples(s, 4)
    dms = np.arange(0, 120, 1.0, dtype=f32)
    plan = _plan(dd, nch, 64e-6, 1581.8, -0.390625)
    plan.set_dm_list(dms)
    nsc = nsamps - plan.max_delay
    d_in = torch.from_numpy(data).cuda()
    d_out = torch.zeros((dms.size, nsc), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    rc = plan.execute_raw(nsamps, d_in, 4, d_out, 32, dd.DEVICE_POINTERS)
    assert rc == 0
    got = d_out.cpu().numpy()
    delays = orc.delay_table(nch, f32(64e-6), f32(1581.8), f32(-0.390625))
    assert np.array_equal(got, orc.tdd(nch, delays, None, dms, nsamps, plan.max_delay, data, 4, 32))


def test_tdd_table_cache_invalidation(dd, orc):
    rng = np.random.default_rng(41)
    nch, nsamps = 64, 2000
    s = rng.integers(0, 256, size=(nsamps, nch)).astype(np.uint32)
    data = dd.pack_samples(s, 8)
    delays = orc.delay_table(nch, f32(64e-6), f32(1581.8), f32(-2.0))
    plan = _plan(dd, nch, 64e-6, 1581.8, -2.0)
    for dms in (np.linspace(0, 10, 5, dtype=f32), np.linspace(3, 25, 9, dtype=f32)):
        plan.set_dm_list(dms)
        for km in (None, (np.arange(nch) % 3 != 0).astype(np.int32)):
            plan.set_killmask(km)
            out = plan.execute(data, 8, 32)
            assert np.array_equal(out, orc.tdd(nch, delays, km, dms, nsamps, plan.max_delay, data, 8, 32))


def test_tdd_cfg2_full_size_properties(dd, orc):
    """1024 ch 1530->1210 MHz, dt 54.6 us, 2^20 4-bit samples, the generated
    1167-trial list: full run on device; DM 0 equals the channel sum and a sample
    of DM rows equals the oracle bit for bit."""
    import torch
    nch, dt, f0, df = 1024, 54.6e-6, 1530.0, -0.3125
    nsamps = 1 << 20
    rng = np.random.default_rng(2)
    s = np.clip(np.rint(rng.normal(8, 2.5, size=(nsamps, nch)).astype(f32)), 0, 15).astype(np.uint8)
    data = dd.pack_samples(s, 4)
    plan = _plan(dd, nch, dt, f0, df)
    plan.generate_dm_list(0.0, 1000.0, 40.0, 1.25)
    assert plan.dm_count == 1167 and plan.max_delay == 19465
    nsc = nsamps - plan.max_delay
    d_in = torch.from_numpy(data).cuda()
    d_out = torch.empty((plan.dm_count, nsc), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    assert plan.execute_raw(nsamps, d_in, 4, d_out, 32, dd.DEVICE_POINTERS) == 0
    dms = plan.dm_list()
    delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    # DM 0: channel sum
    row0 = d_out[0].cpu().numpy()
    sums = s[:nsc].astype(np.int64).sum(axis=1).astype(f32)
    factor = (f32(3.0) * f32(1024.0)) / f32(255.0) / f32(16.0)
    assert np.array_equal(row0, ((sums * f32(1.0)) / (f32(15.0) * f32(nch)) * factor).astype(f32))
    del s
    for i in (1, 583, 1166):
        ref = orc.tdd(nch, delays, None, dms, nsamps, plan.max_delay, data, 4, 32, first_dm=i, dm_count=1)
        assert np.array_equal(d_out[i].cpu().numpy(), ref[0]), i


@pytest.mark.parametrize("nsamps", [1 << 17, 1 << 18])
def test_tdd_cfg4_beams_2bit(dd, orc, nsamps):
    """512 ch 1500 MHz down in 0.78125 MHz, dt 81.92 us, 2-bit N(1.5, 0.8),
    1024 trials 0..2000: the shortest beams of the many-beam workload at full
    geometry. DM 0 equals the channel sum; three DM rows equal the oracle."""
    import torch
    from dedisp_tests_amd import synth
    g = synth.CFG4
    nch, dt, f0, df = g["nchans"], g["dt"], g["f0"], g["df"]
    delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    s = synth.filterbank(nsamps, nch, 2, g["mean"], g["sigma"], seed=nsamps,
                         pulses=[(700.0, 5000, 3, 2)], delays=delays)
    data = dd.pack_samples(s, 2)
    plan = _plan(dd, nch, dt, f0, df)
    dms = np.linspace(0.0, 2000.0, 1024, dtype=f32)
    plan.set_dm_list(dms)
    assert plan.max_delay == 38585
    nsc = nsamps - plan.max_delay
    d_in = torch.from_numpy(data).cuda()
    d_out = torch.empty((1024, nsc), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    assert plan.execute_raw(nsamps, d_in, 2, d_out, 32, dd.DEVICE_POINTERS) == 0
    row0 = d_out[0].cpu().numpy()
    sums = s[:nsc].astype(np.int64).sum(axis=1).astype(f32)
    factor = (f32(3.0) * f32(1024.0)) / f32(255.0) / f32(16.0)
    assert np.array_equal(row0, ((sums * f32(1.0)) / (f32(3.0) * f32(nch)) * factor).astype(f32))
    for i in (358, 511, 1023):  # 358: nearest trial to the injected DM 700
        ref = orc.tdd(nch, delays, None, dms, nsamps, plan.max_delay, data, 2, 32, first_dm=i, dm_count=1)
        assert np.array_equal(d_out[i].cpu().numpy(), ref[0]), i


def test_host_pointers_many_blocks_and_pieces(dd):
    """Host-pointer execute with an output of several DM blocks (about 256 MB
    each) and many 64 MiB transfer pieces, strided input and output rows
    (execute_adv): identical to the device-pointer execute."""
    import torch
    nch, nbits = 64, 8
    plan = dd.Plan(nch, 64e-6, 1581.8, -2.0, impl=dd.TDD)
    plan.set_dm_list(np.linspace(0, 100, 150, dtype=np.float32))
    md, ndm = plan.max_delay, plan.dm_count
    nsamps = (1 << 20) + 12345
    nsc = nsamps - md
    rng = np.random.default_rng(77)
    in_stride, out_stride = nch + 40, nsc * 4 + 256
    host_in = np.zeros((nsamps, in_stride), np.uint8)
    host_in[:, :nch] = rng.integers(0, 256, size=(nsamps, nch), dtype=np.uint8)
    host_out = np.zeros((ndm, out_stride), np.uint8)
    assert plan.execute_adv(nsamps, host_in, nbits, in_stride, host_out, 32, out_stride) == 0
    d_in = torch.from_numpy(np.ascontiguousarray(host_in[:, :nch])).cuda()
    d_out = torch.empty((ndm, nsc), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    assert plan.execute_raw(nsamps, d_in, nbits, d_out, 32, dd.DEVICE_POINTERS) == 0
    torch.cuda.synchronize()
    got = host_out[:, :nsc * 4].copy().view(np.float32)
    assert np.array_equal(got, d_out.cpu().numpy())
    assert not host_out[:, nsc * 4:].any()  # the row padding is not written


def test_host_pointers_rows_longer_than_a_transfer_slot(dd):
    """Output rows of 18 M samples (72 MB, longer than one 64 MiB page-locked
    slot, so each row moves in segments), 1-bit input of 64 channels: the
    host-pointer execute equals the device-pointer one."""
    import torch
    nch, nbits = 64, 1
    plan = dd.Plan(nch, 64e-6, 1581.8, -2.0, impl=dd.TDD)
    plan.set_dm_list(np.array([0.0, 7.0, 30.0], np.float32))
    md, ndm = plan.max_delay, plan.dm_count
    nsamps = 18_000_000
    nsc = nsamps - md
    assert nsc * 4 > (64 << 20)
    rng = np.random.default_rng(3)
    host_in = rng.integers(0, 256, size=(nsamps, nch // 8), dtype=np.uint8)
    out = plan.execute(host_in, nbits, 32)
    d_in = torch.from_numpy(host_in).cuda()
    d_out = torch.empty((ndm, nsc), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    assert plan.execute_raw(nsamps, d_in, nbits, d_out, 32, dd.DEVICE_POINTERS) == 0
    torch.cuda.synchronize()
    assert np.array_equal(out, d_out.cpu().numpy())


@pytest.mark.parametrize("gulp,nbits,out_nbits", [(65536, 4, 32), (10007, 8, 16), (1 << 20, 2, 8)])
def test_tdd_gulped_execution_matches_resident(dd, orc, gulp, nbits, out_nbits):
    """set_gulp_size makes execute run in gulps of that many output samples
    (the reference's TDDPlan.cpp:225-263 loop): outputs are bit-identical to the
    resident single pass, on host and device pointers and with a DM subset
    (execute_guru), the last gulp ragged."""
    import torch
    nch, nsamps = 256, 300001
    rng = np.random.default_rng(gulp)
    s = rng.integers(0, 1 << nbits, size=(nsamps, nch), dtype=np.uint32)
    data = dd.pack_samples(s, nbits)
    delays = orc.delay_table(nch, f32(64e-6), f32(1581.8), f32(-0.390625))
    dms = np.linspace(0, 300, 97, dtype=f32)
    ref_plan = _plan(dd, nch, 64e-6, 1581.8, -0.390625)
    ref_plan.set_dm_list(dms)
    ref = ref_plan.execute(data, nbits, out_nbits)
    plan = _plan(dd, nch, 64e-6, 1581.8, -0.390625)
    plan.set_dm_list(dms)
    assert dd.lib().dedisp_set_gulp_size(plan.handle, gulp) == 0
    assert dd.lib().dedisp_get_gulp_size(plan.handle) == gulp
    got = plan.execute(data, nbits, out_nbits)
    assert np.array_equal(got, ref)
    # device pointers, DMs 10 .. 59 through execute_guru
    nsc = nsamps - plan.max_delay
    tdt = {8: torch.uint8, 16: torch.int16, 32: torch.float32}[out_nbits]
    d_in = torch.from_numpy(data).cuda()
    d_out = torch.zeros((50, nsc), dtype=tdt, device="cuda")
    torch.cuda.synchronize()
    ob = out_nbits // 8
    assert plan.execute_guru(nsamps, d_in, nbits, data.shape[1], d_out, out_nbits, nsc * ob, 10, 50,
                             dd.DEVICE_POINTERS) == 0
    torch.cuda.synchronize()
    got_d = d_out.cpu().numpy().view(ref.dtype)
    assert np.array_equal(got_d, ref[10:60])
    # one row against the oracle as well
    r = orc.tdd(nch, delays, None, dms, nsamps, plan.max_delay, data, nbits, out_nbits, first_dm=96, dm_count=1)
    assert np.array_equal(got[96], r[0])
    assert dd.lib().dedisp_set_gulp_size(plan.handle, 0) == dd.DEDISP_UNKNOWN_ERROR
