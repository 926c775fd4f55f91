"""Parity at the full sizes of the benchmark workloads.

* 1024 ch x 2^20 8-bit samples, 2048 trials 0..1023.5, overlap-save chunks of
  2^16: the three FDD contraction forms against the FP64 oracle on 8 spread DM
  rows and the 5 rows nearest the injected pulses (relative L2 and pulse S/N).
* 64-beam, 512 ch 2-bit workload (1024 trials 0..2000, max delay 38585):
  FDD-BF16 batched execution of 6 of its beams (the longest and the shortest
  among them) against the oracle.
* Streaming workload (1024 ch, 3000 trials 0..1500, 1 s chunks of 15625
  samples, max delay 24849): batched pushes of 2 beams through FDD-BF16 and
  FDD-FP32 against the oracle.
* 1024 ch x 2^20 4-bit samples with the generated 1167-trial list and 5
  injected pulses: TDD bit-exact against the oracle on 32 DM rows, including
  the rows of the pulses.

Inputs are generated on the device (seeded) and copied to the host for the
oracle. Tolerances: relative L2 <= 1e-4 for float32 and the three-product
bfloat16 contraction, <= 1e-3 for the single-product float16 form; pulse S/N
within 0.5 % of the oracle's; TDD bit-exact.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
f32 = np.float32

TOL = {"fp32": 1e-4, "bf16": 1e-4, "fp16": 1e-3}
SNR_TOL = 5e-3


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def snr(x):
    return (x.max() - np.median(x)) / x.std()


def _device_samples(torch, nsamps, nch, nbits, mean, sigma, seed, pulses, delays):
    """Unpacked samples uint8 [nsamps, nch] on the device: clip(round(N(mean,
    sigma))) plus boxcar pulses [(dm, t0, width, amp)] at the TDD delays
    rn(dm * delay[c])."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    vmax = (1 << nbits) - 1
    s = torch.randn((nsamps, nch), generator=g, device="cuda", dtype=torch.float32)
    s = torch.clamp(torch.round(s * sigma + mean), 0, vmax)
    dly = torch.from_numpy(np.asarray(delays, f32)).cuda()
    cols = torch.arange(nch, device="cuda")
    for dm, t0, width, amp in pulses:
        d = torch.round(torch.tensor(f32(dm), device="cuda") * dly).to(torch.int64)
        for w in range(width):
            t = t0 + d + w
            ok = t < nsamps
            s[t[ok], cols[ok]] = torch.clamp(s[t[ok], cols[ok]] + amp, max=vmax)
    return s.to(torch.uint8).contiguous()


def _pack_device(torch, s, nbits):
    """Packs [nsamps, nch] uint8 samples LSB-first along channels on the device."""
    if nbits == 8:
        return s
    per = 8 // nbits
    v = s.view(s.shape[0], s.shape[1] // per, per).to(torch.int32)
    out = torch.zeros(v.shape[:2], dtype=torch.int32, device=s.device)
    for k in range(per):
        out |= v[:, :, k] << (k * nbits)
    return out.to(torch.uint8).contiguous()


def _pulses(seed, n, nsamps, lo, hi, amp):
    rng = np.random.default_rng(seed)
    return [(float(rng.uniform(lo, hi)), int(rng.integers(0, nsamps // 2)), int(rng.integers(1, 17)), amp)
            for _ in range(n)]


# ---------------------------------------------------------------------------
# FDD precision workload
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def fdd_precision_case(dd, orc):
    import torch
    nch, dt, f0, df = 1024, 54.6e-6, 1530.0, -0.3125
    nsamps = 1 << 20
    delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    pulses = _pulses(31, 5, nsamps, 50.0, 900.0, 30)
    d_in = _device_samples(torch, nsamps, nch, 8, 128.0, 20.0, 31, pulses, delays)
    host = d_in.cpu().numpy()
    dms = (np.arange(2048, dtype=f32) * f32(0.5)).astype(f32)
    rows = [0, 1, 293, 586, 1024, 1317, 1755, 2047]
    prow = [int(round(p[0] / 0.5)) for p in pulses]
    sel = rows + prow
    # oracle with max delay 0: its first nsamps - max_delay outputs are the
    # plan's (the shift-sum reads zeros beyond the series)
    ref = orc.fdd_timeseries(nch, delays, dms[sel], nsamps, 0, host)
    del host
    return dict(nch=nch, dt=dt, f0=f0, df=df, nsamps=nsamps, d_in=d_in, dms=dms, sel=sel,
                nrows=len(rows), ref=ref)


@pytest.mark.parametrize("form", ["fp32", "bf16", "fp16"])
def test_fdd_precision_workload_forms(dd, fdd_precision_case, form):
    import torch
    c = fdd_precision_case
    prec = {"fp32": dd.FDD_FP32, "bf16": dd.FDD_BF16, "fp16": dd.FDD_FP16}[form]
    plan = dd.Plan(c["nch"], c["dt"], c["f0"], c["df"], impl=dd.FDD)
    plan.fdd_set_output(dd.FDD_TIMESERIES)
    plan.fdd_set_segmented(True, 65536)
    plan.fdd_set_precision(prec)
    plan.set_dm_list(c["dms"])
    assert plan.max_delay == 19874
    nsc = c["nsamps"] - plan.max_delay
    d_out = torch.empty((2048, nsc), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    print(f"fdd precision workload: form {form}", flush=True)
    assert plan.execute_raw(c["nsamps"], c["d_in"], 8, d_out, 32, dd.DEVICE_POINTERS) == 0
    torch.cuda.synchronize()
    got = d_out[c["sel"]].cpu().numpy()
    del d_out
    ref = c["ref"][:, :nsc]
    n = c["nrows"]
    err = rel_l2(got[:n], ref[:n])
    assert err < TOL[form], err
    for j in range(n, len(c["sel"])):
        assert rel_l2(got[j], ref[j]) < TOL[form]
        assert abs(snr(got[j]) / snr(ref[j]) - 1) < SNR_TOL, (form, c["sel"][j])
    plan.close()


# ---------------------------------------------------------------------------
# Many-beam workload: FDD-BF16 batched execution
# ---------------------------------------------------------------------------

def test_many_beam_workload_fdd_bf16_batch(dd, orc):
    import torch
    from dedisp_tests_amd import synth
    g = synth.CFG4
    nch, dt, f0, df = g["nchans"], g["dt"], g["f0"], g["df"]
    lo, hi = g["nsamps_range"]
    rng = np.random.default_rng(2024)
    lengths = np.exp(rng.uniform(np.log(lo), np.log(hi), size=g["beams"])).astype(np.int64)
    order = np.argsort(lengths)
    pick = [int(order[-1]), int(order[0]), int(order[10]), int(order[25]), int(order[40]), int(order[55])]
    delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    dms = np.linspace(0.0, 2000.0, 1024, dtype=f32)
    plan = dd.Plan(nch, dt, f0, df, impl=dd.FDD)
    plan.fdd_set_output(dd.FDD_TIMESERIES)
    plan.fdd_set_precision(dd.FDD_BF16)
    plan.set_dm_list(dms)
    assert plan.max_delay == 38585
    md = plan.max_delay
    ins, outs, hosts, ns = [], [], [], []
    for k, b in enumerate(pick):
        n = int(lengths[b])
        pul = [(float(p[0]), p[1], p[2], 2) for p in _pulses(700 + b, 1 + k % 3, n, 100.0, 1900.0, 2)]
        s = _device_samples(torch, n, nch, 2, g["mean"], g["sigma"], 700 + b, pul, delays)
        hosts.append(s.cpu().numpy())
        ins.append(_pack_device(torch, s, 2))
        del s
        outs.append(torch.empty((1024, n - md), dtype=torch.float32, device="cuda"))
        ns.append(n)
    torch.cuda.synchronize()
    print(f"many-beam FDD-BF16 batch: lengths {ns}", flush=True)
    assert plan.execute_batch_raw(ns, ins, 2, outs, 32, dd.DEVICE_POINTERS) == 0
    torch.cuda.synchronize()
    rows = [0, 1, 300, 511, 700, 1023]
    for k in range(len(pick)):
        ref = orc.fdd_timeseries(nch, delays, dms[rows], ns[k], md, hosts[k], nbits=2)
        got = outs[k][rows].cpu().numpy()
        assert rel_l2(got, ref) < TOL["bf16"], (k, ns[k])
    plan.close()


# ---------------------------------------------------------------------------
# Streaming workload: batched pushes, FDD-BF16 and FDD-FP32
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("form", ["bf16", "fp32"])
def test_streaming_workload_fdd_batched_pushes(dd, orc, form):
    import torch
    from dedisp_tests_amd import synth
    g = synth.CFG5
    nch, dt, f0, df, chunk = g["nchans"], g["dt"], g["f0"], g["df"], g["chunk"]
    plan = dd.Plan(nch, dt, f0, df, impl=dd.FDD)
    plan.fdd_set_precision(dd.FDD_BF16 if form == "bf16" else dd.FDD_FP32)
    dms = np.linspace(0.0, 1500.0, 3000, dtype=f32)
    plan.set_dm_list(dms)
    hist = plan.max_delay
    assert hist == 24849
    nb, npush = 2, 3
    delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    st = dd.Stream(plan, nbeams=nb, chunk_nsamps=chunk, in_nbits=8, out_nbits=32)
    series = []
    for b in range(nb):
        pul = _pulses(90 + b, 2, chunk * npush, 100.0, 1400.0, 30)
        series.append(_device_samples(torch, chunk * npush, nch, 8, 128.0, 20.0, 90 + b, pul, delays))
    d_out = [torch.empty((3000, chunk), dtype=torch.float32, device="cuda") for _ in range(nb)]
    got = [[] for _ in range(nb)]
    torch.cuda.synchronize()
    print(f"streaming FDD batched pushes: form {form}", flush=True)
    for n in range(npush):
        chunks = [series[b][n * chunk:(n + 1) * chunk] for b in range(nb)]
        assert st.push_batch_raw(list(range(nb)), chunks, d_out, dd.DEVICE_POINTERS) == 0
        torch.cuda.synchronize()
        for b in range(nb):
            got[b].append(d_out[b].cpu().numpy().copy())
    rows = [0, 1, 750, 1499, 2250, 2999]
    for b in range(nb):
        full = np.concatenate(got[b], axis=1)[:, hist:]        # outputs t >= 0
        host = series[b].cpu().numpy()
        ref = orc.fdd_timeseries(nch, delays, dms[rows], chunk * npush, hist, host)
        assert full.shape[1] == ref.shape[1]
        assert rel_l2(full[rows], ref) < TOL[form], (form, b)
    st.close()
    plan.close()


# ---------------------------------------------------------------------------
# Large TDD workload: 32 DM rows bit-exact, pulses recovered
# ---------------------------------------------------------------------------

def test_tdd_large_workload_32_rows(dd, orc):
    import torch
    nch, dt, f0, df = 1024, 54.6e-6, 1530.0, -0.3125
    nsamps = 1 << 20
    delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    plan = dd.Plan(nch, dt, f0, df, impl=dd.TDD)
    plan.generate_dm_list(0.0, 1000.0, 40.0, 1.25)
    dms = plan.dm_list()
    assert plan.dm_count == 1167 and plan.max_delay == 19465
    # pulses at trial DMs, so the matching trial sums each one exactly
    prng = np.random.default_rng(77)
    pulses = []
    for _ in range(5):
        i = int(np.argmin(np.abs(dms - prng.uniform(50, 900))))
        pulses.append((float(dms[i]), int(prng.integers(0, nsamps // 2)), int(prng.integers(1, 17)), 4))
    s = _device_samples(torch, nsamps, nch, 4, 8.0, 2.5, 77, pulses, delays)
    d_in = _pack_device(torch, s, 4)
    host = d_in.cpu().numpy()
    del s
    nsc = nsamps - plan.max_delay
    d_out = torch.empty((plan.dm_count, nsc), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    print("tdd large workload: 1167 trials", flush=True)
    assert plan.execute_raw(nsamps, d_in, 4, d_out, 32, dd.DEVICE_POINTERS) == 0
    torch.cuda.synchronize()
    prow = [int(np.argmin(np.abs(dms - p[0]))) for p in pulses]
    rows = sorted(set(np.linspace(0, 1166, 27).astype(int).tolist() + prow))
    assert len(rows) >= 32
    got = d_out[rows].cpu().numpy()
    # the oracle over the selected trials in one call (its DM loop runs in parallel)
    ref = orc.tdd(nch, delays, None, dms[rows], nsamps, plan.max_delay, host, 4, 32)
    for j, i in enumerate(rows):
        assert np.array_equal(got[j], ref[j]), i
    # every pulse peaks at its own trial, at its start sample
    for (dm, t0, width, amp), i in zip(pulses, prow):
        row = got[rows.index(i)]
        assert t0 <= int(np.argmax(row)) < t0 + width, (dm, t0)
    plan.close()
