"""Filterbank front end: SIGPROC reading, decimation, PRESTO .dat/.inf writing
(dedisp_amd_fil_*, dedisp_amd_write_*) and the dedisp_fil tool.

The files are written here by a small SIGPROC writer (length-prefixed keywords,
little-endian values), including keywords the reader does not interpret, whose
values the reader has to step over.
"""
import ctypes as C
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

f32 = np.float32


def _kw(name: str) -> bytes:
    b = name.encode()
    return struct.pack("<i", len(b)) + b


def write_fil(path: Path, data: np.ndarray, nbits=8, tsamp=64e-6, fch1=1581.8, foff=-0.78125,
              tstart=59000.123456789012, source="J0000+0000", raj=123456.789, dej=-123456.7,
              telescope_id=4, machine_id=10, extra=True):
    nsamps, row = data.shape
    nchans = row * 8 // nbits
    h = _kw("HEADER_START")
    if extra:  # keywords the reader steps over
        h += _kw("rawdatafile") + _kw("some_raw_file.raw")
        h += _kw("data_type") + struct.pack("<i", 1)
        h += _kw("refdm") + struct.pack("<d", 0.0)
    h += _kw("source_name") + _kw(source)
    h += _kw("telescope_id") + struct.pack("<i", telescope_id)
    h += _kw("machine_id") + struct.pack("<i", machine_id)
    h += _kw("src_raj") + struct.pack("<d", raj)
    h += _kw("src_dej") + struct.pack("<d", dej)
    h += _kw("tstart") + struct.pack("<d", tstart)
    h += _kw("tsamp") + struct.pack("<d", tsamp)
    h += _kw("fch1") + struct.pack("<d", fch1)
    h += _kw("foff") + struct.pack("<d", foff)
    h += _kw("nchans") + struct.pack("<i", nchans)
    h += _kw("nbits") + struct.pack("<i", nbits)
    h += _kw("nifs") + struct.pack("<i", 1)
    h += _kw("HEADER_END")
    path.write_bytes(h + np.ascontiguousarray(data, dtype=np.uint8).tobytes())
    return len(h)


def test_read_header_and_data(dd, tmp_path):
    rng = np.random.default_rng(1)
    data = rng.integers(0, 256, size=(1000, 96), dtype=np.uint8)
    f = tmp_path / "a.fil"
    hbytes = write_fil(f, data)
    h, got = dd.read_filterbank(f)
    d = h.as_dict()
    assert d["header_bytes"] == hbytes and d["data_bytes"] == data.nbytes
    assert (d["nchans"], d["nsamps"], d["nbits"], d["nifs"]) == (96, 1000, 8, 1)
    assert d["tsamp"] == 64e-6 and d["fch1"] == 1581.8 and d["foff"] == -0.78125
    assert d["tstart"] == 59000.123456789012 and d["src_dej"] == -123456.7
    assert (d["telescope_id"], d["machine_id"]) == (4, 10)
    assert d["source_name"] == "J0000+0000"
    assert np.array_equal(got, data)


def test_read_header_errors(dd, tmp_path):
    h = dd.FilHeader()
    assert dd.lib().dedisp_amd_fil_read_header(str(tmp_path / "missing.fil").encode(), C.byref(h)) == \
        dd.DEDISP_UNKNOWN_ERROR
    f = tmp_path / "trunc.fil"
    f.write_bytes(_kw("HEADER_START") + _kw("nchans"))  # ends inside the header
    assert dd.lib().dedisp_amd_fil_read_header(str(f).encode(), C.byref(h)) == dd.DEDISP_UNKNOWN_ERROR


@pytest.mark.parametrize("factor", [1, 2, 3, 8])
def test_decimate_matches_numpy(dd, factor):
    rng = np.random.default_rng(factor)
    nch, ns = 40, 203
    data = rng.integers(0, 256, size=(ns, nch), dtype=np.uint8)
    buf = data.copy()
    n = dd.lib().dedisp_amd_fil_decimate(dd._ptr(buf), nch, ns, factor)
    assert n == ns // factor
    # float32 running sum over the group, divided by the factor, truncated
    g = data[: n * factor].reshape(n, factor, nch).astype(np.float32)
    acc = np.zeros((n, nch), np.float32)
    for k in range(factor):
        acc = (acc + g[:, k]).astype(np.float32)
    want = np.clip(np.trunc(acc / np.float32(factor)), 0, 255).astype(np.uint8)
    assert np.array_equal(buf[:n], want)


def test_write_dat_and_inf(dd, tmp_path):
    data = np.zeros((10, 64), np.uint8)
    f = tmp_path / "b.fil"
    write_fil(f, data, tsamp=6.4e-05, fch1=1400.0, foff=-0.5, tstart=58000.25, source="B1937+21",
              raj=193939.1, dej=213459.8, telescope_id=11, machine_id=11)
    h, _ = dd.read_filterbank(f)
    prefix = str(tmp_path / "out")
    series = np.arange(6, dtype=np.float32)
    assert dd.lib().dedisp_amd_write_dat(prefix.encode(), f32(12.5), dd._ptr(series), 6) == 0
    assert np.array_equal(np.fromfile(prefix + "_DM12.500.dat", dtype=np.float32), series)
    assert dd.lib().dedisp_amd_write_inf(prefix.encode(), C.byref(h), f32(12.5), 6) == 0
    text = Path(prefix + "_DM12.500.inf").read_text().splitlines()
    fields = {l.split("=")[0].strip(): l.split("=", 1)[1].strip() for l in text if "=" in l}
    assert fields["Data file name without suffix"] == prefix + "_DM12.500"
    assert fields["Telescope used"] == "LOFAR"
    assert fields["Instrument used"] == "Cobalt"
    assert fields["Object being observed"] == "B1937+21"
    assert fields["J2000 Right Ascension (hh:mm:ss.ssss)"] == "19:39:39.100000"
    assert fields["J2000 Declination     (dd:mm:ss.ssss)"] == "21:34:59.800000"
    assert fields["Epoch of observation (MJD)"] == "58000.250000000000000"
    assert fields["Number of bins in the time series"] == "6"
    assert fields["Width of each time series bin (sec)"] == "6.4e-05"
    assert fields["Dispersion measure (cm-3 pc)"] == "12.5"
    assert fields["Central freq of low channel (Mhz)"] == "1368.5"
    assert fields["Total bandwidth (Mhz)"] == "32"
    assert fields["Number of channels"] == "64"
    assert fields["Channel bandwidth (Mhz)"] == "0.5"
    assert text[-1].strip() == "Input filterbank samples have 8 bits."
    # negative declination keeps the sign in front of the degrees
    write_fil(tmp_path / "a2.fil", data, dej=-51234.5)
    h2, _ = dd.read_filterbank(tmp_path / "a2.fil")
    assert dd.lib().dedisp_amd_write_inf((prefix + "2").encode(), C.byref(h2), f32(0.0), 6) == 0
    dec = [l for l in Path(prefix + "2_DM0.000.inf").read_text().splitlines() if "Declination" in l][0]
    assert dec.split("=", 1)[1] == " -05:12:34.500000"


def test_cli_usage_runs_without_device(dd):
    exe = Path(dd.__file__).parent / "bin" / "dedisp_fil"
    assert exe.exists(), "build the library first"
    r = subprocess.run([str(exe), "-h"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "SIGPROC filterbank" in r.stdout


@pytest.mark.gpu
def test_cli_end_to_end_matches_oracle(dd, orc, tmp_path):
    nch, ns = 128, 12000
    dt, f0, df = 64e-6, 1581.8, -0.78125
    rng = np.random.default_rng(21)
    data = rng.integers(0, 256, size=(ns, nch), dtype=np.uint8)
    f = tmp_path / "c.fil"
    write_fil(f, data, tsamp=dt, fch1=f0, foff=df)
    exe = Path(dd.__file__).parent / "bin" / "dedisp_fil"
    prefix = str(tmp_path / "trial")
    r = subprocess.run([str(exe), "-f", str(f), "-o", prefix, "-i", "tdd", "-r", "10", "-s", "7.5", "-n", "5",
                        "-N", "9001", "-q"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    dms = (f32(10) + f32(7.5) * np.arange(5, dtype=f32)).astype(f32)
    delays = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    md = orc.max_delay(float(dms[-1]), float(delays[-1]))
    ref = orc.tdd(nch, delays, np.ones(nch, np.int32), dms, ns, md, data, 8, 32)
    for i, dm in enumerate(dms):
        got = np.fromfile(f"{prefix}_DM{dm:.3f}.dat", dtype=np.float32)
        assert got.size == 9000  # -N 9001 made even
        assert np.array_equal(got, ref[i, :9000])
        assert Path(f"{prefix}_DM{dm:.3f}.inf").exists()
