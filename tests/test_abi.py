"""CPU tests of the C-ABI library: it loads, exports every declared symbol, and
its host-only logic (error strings, delay table, DM list, null-plan handling)
matches the reference behaviour. No compute call is made here."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import gpu_available

REPO = Path(__file__).resolve().parents[1]
f32 = np.float32


def _declared_functions():
    names = set()
    for h in (REPO / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(dedisp_\w+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol(dd):
    out = subprocess.run(["nm", "-D", "--defined-only", str(dd.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    declared = _declared_functions()
    assert len(declared) >= 38
    missing = [n for n in declared if n not in exported]
    assert not missing, missing
    # the Python binding covers the same set
    assert set(dd.EXPORTED) == set(declared)


def test_library_needs_no_oracle(dd):
    out = subprocess.run(["ldd", str(dd.LIB_PATH)], capture_output=True, text=True).stdout
    assert "oracle" not in out
    syms = subprocess.run(["nm", "-D", str(dd.LIB_PATH)], capture_output=True, text=True).stdout
    assert "oracle_" not in syms


def test_library_has_gfx950_code(dd):
    data = dd.LIB_PATH.read_bytes()
    assert b"gfx950" in data


def test_peak_probe_library_is_separate(dd):
    # the bench's integer-add probe ships in its own library, not in libdedisp.so
    peaks = dd.LIB_PATH.parent / "libdd_peaks.so"
    out = subprocess.run(["nm", "-D", "--defined-only", str(peaks)], capture_output=True,
                         text=True, check=True).stdout
    assert "dd_peak_int_add" in out
    assert b"gfx950" in peaks.read_bytes()
    syms = subprocess.run(["nm", "-D", str(dd.LIB_PATH)], capture_output=True, text=True).stdout
    assert "dd_peak_" not in syms


def test_error_strings(dd):
    # src/common/dedisp_error.cpp:3-42
    assert dd.error_string(dd.DEDISP_NO_ERROR) == "No error"
    assert dd.error_string(dd.DEDISP_TOO_FEW_NSAMPS) == "No. samples < maximum delay"
    assert dd.error_string(dd.DEDISP_UNSUPPORTED_IN_NBITS) == "Unsupported in_nbits value"
    assert dd.error_string(999) == "Invalid error code"
    assert len(dd.ERRORS) == 17


def test_null_plan_errors(dd):
    l = dd.lib()
    assert l.dedisp_set_dm_list(None, None, 0) == dd.DEDISP_INVALID_PLAN
    assert l.dedisp_execute(None, 0, None, 8, None, 32, 0) == dd.DEDISP_INVALID_PLAN
    assert l.dedisp_execute_adv(None, 0, None, 8, 0, None, 32, 0, 0) == dd.DEDISP_INVALID_PLAN
    assert l.dedisp_execute_guru(None, 0, None, 8, 0, None, 32, 0, 0, 0, 0) == dd.DEDISP_INVALID_PLAN
    assert l.dedisp_set_gulp_size(None, 10) == dd.DEDISP_INVALID_PLAN
    assert l.dedisp_get_dm_count(None) == 0
    assert l.dedisp_get_max_delay(None) == 0
    assert not l.dedisp_get_dm_list(None)
    assert l.dedisp_create_plan(None, 16, 1e-4, 1400.0, -1.0) == dd.DEDISP_INVALID_POINTER
    assert l.dedisp_select_implementation(7) == dd.DEDISP_UNKNOWN_ERROR


@pytest.mark.parametrize("geom", [(256, 64e-6, 1581.8, -0.390625), (1024, 54.6e-6, 1530.0, -0.3125),
                                  (512, 81.92e-6, 1500.0, -0.78125), (100, 1e-3, 400.0, 2.0)])
def test_product_delay_table_matches_oracle(dd, orc, geom):
    nch, dt, f0, df = geom
    a = dd.generate_delay_table(nch, dt, f0, df)
    b = orc.delay_table(nch, f32(dt), f32(f0), f32(df))
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_product_dm_list_matches_golden(dd):
    g = np.load(REPO / "tests" / "golden" / "plan_tables.npz")
    dml = dd.generate_dm_list_guru(0.0, 1000.0, float(f32(54.6e-6)), 40.0, float(f32(1530.0)),
                                   float(f32(-0.3125)), 1024, 1.25)
    assert np.array_equal(dml.view(np.uint32), g["cfg2_dm_list_bits"])


def test_product_dm_list_terminates_when_not_advancing(dd):
    # tol = 1 with ti = dt = 0 gives a recurrence that never advances; the
    # reference loops forever, this library stops.
    dml = dd.generate_dm_list_guru(1.0, 10.0, 0.0, 0.0, 1400.0, -1.0, 64, 1.0)
    assert dml.size >= 1 and dml[0] == 1.0


@pytest.mark.skipif(gpu_available(), reason="checks the no-device error path")
def test_create_plan_without_device_fails_cleanly(dd):
    import ctypes as C
    h = C.c_void_p(123)
    rc = dd.lib().dedisp_create_plan(C.byref(h), 64, 1e-4, 1400.0, -1.0)
    assert rc != dd.DEDISP_NO_ERROR
    assert not h.value


def test_use_cpu_fdd_is_refused(dd, monkeypatch):
    import ctypes as C
    monkeypatch.setenv("USE_CPU", "1")
    dd.select_implementation(dd.FDD)
    try:
        h = C.c_void_p(0)
        assert dd.lib().dedisp_create_plan(C.byref(h), 64, 1e-4, 1400.0, -1.0) == dd.DEDISP_UNKNOWN_ERROR
        assert not h.value
    finally:
        dd.select_implementation(dd.DEDISP)


def test_pack_unpack_roundtrip(dd):
    rng = np.random.default_rng(0)
    for nbits in (1, 2, 4, 8, 16, 32):
        hi = (1 << nbits) if nbits < 32 else (1 << 32)
        s = rng.integers(0, hi, size=(17, 64), dtype=np.uint64).astype(np.uint32)
        p = dd.pack_samples(s, nbits)
        assert p.shape == (17, 64 * nbits // 8)
        assert np.array_equal(dd.unpack_samples(p, nbits, 64), s)
    # channel 0 in the low bits of byte 0 (LSB first)
    s = np.zeros((1, 8), np.uint32)
    s[0, 0] = 0x5
    s[0, 1] = 0xA
    assert dd.pack_samples(s, 4)[0, 0] == 0xA5
