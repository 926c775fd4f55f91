"""Allocation-free execution (dedisp_amd_execute_noalloc) against dedisp_execute.

The entry point runs the TDD path on device pointers with its rows and delay
tables in the library's fixed device scratch. Its outputs must equal those of
dedisp_execute (device pointers) bit for bit, on consecutive calls whose
second input is smaller (the scratch keeps the first call's contents), across
plans that take turns in the shared scratch (tables re-uploaded), and on the
weighted-killmask path; sizes beyond the scratch are refused before any
device work.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
f32 = np.float32


def _device_input(torch, rng, nsamps, nch, nbits):
    from dedisp_tests_amd import pack_samples
    s = rng.integers(0, 1 << nbits, size=(nsamps, nch), dtype=np.uint32)
    return torch.from_numpy(pack_samples(s, nbits)).cuda()


def _out(torch, plan, nsamps, out_nbits, fill):
    dt = {8: torch.uint8, 16: torch.int16, 32: torch.int32}[out_nbits]
    return torch.full((plan.dm_count, nsamps - plan.max_delay), fill, dtype=dt, device="cuda")


def _reference(torch, dd, plan, d_in, nsamps, nbits, out_nbits):
    ref = _out(torch, plan, nsamps, out_nbits, 0)
    torch.cuda.synchronize()
    assert plan.execute_raw(nsamps, d_in, nbits, ref, out_nbits, dd.DEVICE_POINTERS) == 0
    return ref


def test_noalloc_equals_execute_two_calls_second_smaller(dd):
    import torch
    nch, nbits = 512, 4
    plan = dd.Plan(nch, 64e-6, 1581.8, -0.390625, impl=dd.TDD)
    plan.set_dm_list(np.linspace(0, 400, 150, dtype=f32))
    rng = np.random.default_rng(41)
    sizes = (200_000, 60_001)              # second call smaller, ragged length
    ins = [_device_input(torch, rng, n, nch, nbits) for n in sizes]
    outs = [_out(torch, plan, n, 32, -1) for n in sizes]
    torch.cuda.synchronize()
    for n, d_in, o in zip(sizes, ins, outs):     # two noalloc calls in a row
        print(f"noalloc: {nch} ch {nbits}-bit nsamps {n}", flush=True)
        assert plan.execute_noalloc_raw(n, d_in, nbits, o, 32, 0) == 0
    torch.cuda.synchronize()
    for n, d_in, o in zip(sizes, ins, outs):
        ref = _reference(torch, dd, plan, d_in, n, nbits, 32)
        torch.cuda.synchronize()
        assert torch.equal(o, ref), n
    plan.close()


def test_noalloc_plans_take_turns_in_the_scratch(dd):
    """Two plans (a 0/1-killmask 8-bit plan and a weighted-killmask 2-bit plan,
    which runs the generic kernel) alternate: each call must find its own
    tables, re-uploaded when the other plan's are in the scratch."""
    import torch
    rng = np.random.default_rng(7)
    a = dd.Plan(256, 64e-6, 1581.8, -0.390625, impl=dd.TDD)
    a.set_dm_list(np.arange(0, 120, 1.5, dtype=f32))
    b = dd.Plan(128, 64e-6, 1400.0, -1.0, impl=dd.TDD)
    b.set_dm_list(np.linspace(10, 90, 41, dtype=f32))
    km = np.ones(128, np.int32)
    km[::7] = 0
    km[3::11] = 2
    b.set_killmask(km)
    cases = [(a, 8, 8, 50_000), (b, 2, 16, 40_000), (a, 8, 8, 30_000), (b, 2, 16, 9_000)]
    for plan, nbits, out_nbits, n in cases:
        d_in = _device_input(torch, rng, n, plan.channel_count, nbits)
        o = _out(torch, plan, n, out_nbits, -1)
        torch.cuda.synchronize()
        print(f"noalloc: {plan.channel_count} ch {nbits}-bit out {out_nbits} nsamps {n}", flush=True)
        assert plan.execute_noalloc_raw(n, d_in, nbits, o, out_nbits, dd.DEVICE_POINTERS) == 0
        torch.cuda.synchronize()
        ref = _reference(torch, dd, plan, d_in, n, nbits, out_nbits)
        torch.cuda.synchronize()
        assert torch.equal(o, ref), (plan.channel_count, n)
    a.close()
    b.close()


def test_noalloc_allocates_no_device_memory(dd):
    """A plan that has never executed makes its first call through the
    allocation-free entry point: the device's free memory is unchanged,
    although its rows alone take about 64 MiB."""
    import torch
    nch, nbits, n = 1024, 8, 65_536
    warm = dd.Plan(64, 64e-6, 1581.8, -2.0, impl=dd.TDD)   # loads the code object
    warm.set_dm_list(np.array([0.0, 5.0], f32))
    rng = np.random.default_rng(3)
    w_in = _device_input(torch, rng, 4096, 64, 8)
    w_out = _out(torch, warm, 4096, 32, 0)
    plan = dd.Plan(nch, 64e-6, 1530.0, -0.3125, impl=dd.TDD)
    plan.set_dm_list(np.linspace(0, 50, 64, dtype=f32))
    d_in = _device_input(torch, rng, n, nch, nbits)
    o = _out(torch, plan, n, 32, -1)
    torch.cuda.synchronize()
    assert warm.execute_noalloc_raw(4096, w_in, 8, w_out, 32) == 0
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    assert plan.execute_noalloc_raw(n, d_in, nbits, o, 32) == 0
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert abs(free0 - free1) < (8 << 20), (free0, free1)
    ref = _reference(torch, dd, plan, d_in, n, nbits, 32)
    torch.cuda.synchronize()
    assert torch.equal(o, ref)
    plan.close()
    warm.close()


def test_noalloc_refusals(dd):
    """Rows beyond the fixed scratch, host pointers and FDD plans are refused
    with an error code before any device work; the scratch stays usable."""
    import torch
    nch = 1024
    plan = dd.Plan(nch, 64e-6, 1530.0, -0.3125, impl=dd.TDD)
    plan.set_dm_list(np.array([0.0, 2.0], f32))
    n = 140_000                             # 8-bit rows: 1024 x 140000 bytes > 128 MiB
    assert nch * n > dd.NOALLOC_SCRATCH_BYTES
    d_in = torch.zeros((n, nch), dtype=torch.uint8, device="cuda")
    o = _out(torch, plan, n, 32, 0)
    torch.cuda.synchronize()
    assert plan.execute_noalloc_raw(n, d_in, 8, o, 32) == dd.DEDISP_MEM_ALLOC_FAILED
    assert plan.execute_noalloc_raw(n, d_in, 8, o, 32, dd.HOST_POINTERS) == dd.DEDISP_INVALID_FLAG_COMBINATION
    assert plan.execute_noalloc_raw(n, d_in, 3, o, 32) == dd.DEDISP_UNSUPPORTED_IN_NBITS
    assert plan.execute_noalloc_raw(n, None, 8, o, 32) == dd.DEDISP_INVALID_POINTER
    f = dd.Plan(64, 64e-6, 1581.8, -2.0, impl=dd.FDD)
    f.set_dm_list(np.array([0.0, 5.0], f32))
    assert f.execute_noalloc_raw(4096, d_in, 8, o, 32) == dd.DEDISP_INVALID_PLAN
    # a fitting call afterwards still matches dedisp_execute
    m = 20_000
    o2 = _out(torch, plan, m, 32, -1)
    torch.cuda.synchronize()
    assert plan.execute_noalloc_raw(m, d_in, 8, o2, 32) == 0
    torch.cuda.synchronize()
    ref = _reference(torch, dd, plan, d_in, m, 8, 32)
    torch.cuda.synchronize()
    assert torch.equal(o2, ref)
    f.close()
    plan.close()
