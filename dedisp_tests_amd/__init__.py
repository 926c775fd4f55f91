"""MI355X-native dedispersion: Python binding of libdedisp.so.

The library is C/C++/HIP (dedisp_tests_amd/csrc, public headers in include/);
this module binds its C ABI with ctypes so Python callers and the tests can use
the same entry points a C caller links against (reference src/clib/dedisp.h).

    import dedisp_tests_amd as dd
    plan = dd.Plan(nchans=256, dt=64e-6, f0=1581.8, df=-0.390625, impl=dd.TDD)
    plan.set_dm_list(np.arange(256, dtype=np.float32))
    out = plan.execute(data, in_nbits=8, out_nbits=32)   # (ndm, nsamps - max_delay)

Importing never loads the oracle; the library must exist (build it with
``python -m dedisp_tests_amd.build``) or loading raises.
"""
from __future__ import annotations

import ctypes as C
import weakref
import os
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libdedisp.so"

# dedisp_error (include/dedisp.h)
ERRORS = [
    "DEDISP_NO_ERROR", "DEDISP_MEM_ALLOC_FAILED", "DEDISP_MEM_COPY_FAILED",
    "DEDISP_NCHANS_EXCEEDS_LIMIT", "DEDISP_INVALID_PLAN", "DEDISP_INVALID_POINTER",
    "DEDISP_INVALID_STRIDE", "DEDISP_NO_DM_LIST_SET", "DEDISP_TOO_FEW_NSAMPS",
    "DEDISP_INVALID_FLAG_COMBINATION", "DEDISP_UNSUPPORTED_IN_NBITS",
    "DEDISP_UNSUPPORTED_OUT_NBITS", "DEDISP_INVALID_DEVICE_INDEX", "DEDISP_DEVICE_ALREADY_SET",
    "DEDISP_PRIOR_GPU_ERROR", "DEDISP_INTERNAL_GPU_ERROR", "DEDISP_UNKNOWN_ERROR",
]
for _i, _n in enumerate(ERRORS):
    globals()[_n] = _i

# dedisp_implementation
DEDISP, TDD, FDD = 0, 1, 2
# dedisp_flag
USE_DEFAULT, HOST_POINTERS, DEVICE_POINTERS, WAIT, ASYNC = 0, 1 << 1, 1 << 2, 1 << 3, 1 << 4
# dedisp_amd_fdd_output / precision
FDD_FOURIER, FDD_TIMESERIES = 0, 1
# bytes of the fixed device scratch behind dedisp_amd_execute_noalloc (dedisp_amd.h)
NOALLOC_SCRATCH_BYTES = 128 << 20
FDD_FP32, FDD_BF16, FDD_FP16 = 0, 1, 2


class DedispError(RuntimeError):
    def __init__(self, code: int, where: str):
        self.code = code
        name = ERRORS[code] if 0 <= code < len(ERRORS) else str(code)
        super().__init__(f"{where}: {name} ({error_string(code)})")


class Profile(C.Structure):
    _fields_ = [
        ("h2d_ms", C.c_double), ("unpack_ms", C.c_double), ("unpack_launches", C.c_ulong),
        ("fft_ms", C.c_double), ("fft_launches", C.c_ulong),
        ("dedisperse_ms", C.c_double), ("dedisperse_launches", C.c_ulong),
        ("ifft_ms", C.c_double), ("ifft_launches", C.c_ulong),
        ("d2h_ms", C.c_double), ("total_ms", C.c_double), ("work_units", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class StreamStats(C.Structure):
    _fields_ = [
        ("pushes", C.c_ulong), ("mean_ms", C.c_double), ("p50_ms", C.c_double),
        ("p99_ms", C.c_double), ("max_ms", C.c_double), ("samples_per_s", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class Tuning(C.Structure):
    """dedisp_amd_tuning (include/dedisp_amd.h)."""
    _fields_ = [("samples_per_lane", C.c_uint), ("dms_per_wave", C.c_uint), ("waves", C.c_uint),
                ("chunk_channels", C.c_uint), ("nfft", C.c_ulong), ("best_ms", C.c_double),
                ("candidates", C.c_uint), ("fdd_form", C.c_int), ("fdd_mirror", C.c_int),
                ("fdd_layout", C.c_int)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class FilHeader(C.Structure):
    _fields_ = [
        ("header_bytes", C.c_long), ("data_bytes", C.c_long),
        ("nchans", C.c_int), ("nsamps", C.c_int), ("nbits", C.c_int), ("nifs", C.c_int),
        ("machine_id", C.c_int), ("telescope_id", C.c_int), ("nbeams", C.c_int), ("ibeam", C.c_int),
        ("tstart", C.c_double), ("tsamp", C.c_double), ("fch1", C.c_double), ("foff", C.c_double),
        ("src_raj", C.c_double), ("src_dej", C.c_double), ("az_start", C.c_double), ("za_start", C.c_double),
        ("source_name", C.c_char * 80),
    ]

    def as_dict(self) -> dict:
        d = {n: getattr(self, n) for n, _ in self._fields_}
        d["source_name"] = self.source_name.decode()
        return d


_lib = None

# (name, restype, argtypes) of every exported entry point of include/*.h
_SIG = [
    ("dedisp_amd_execute_batch", C.c_int, [C.c_void_p, C.c_ulong, C.POINTER(C.c_ulong), C.POINTER(C.c_void_p),
                                           C.c_ulong, C.POINTER(C.c_void_p), C.c_ulong, C.c_uint]),
    ("dedisp_amd_autotune", C.c_int, [C.c_void_p, C.c_ulong, C.c_ulong, C.c_ulong, C.c_void_p]),
    ("dedisp_amd_execute_noalloc", C.c_int,
     [C.c_void_p, C.c_ulong, C.c_void_p, C.c_ulong, C.c_void_p, C.c_ulong, C.c_uint]),
    ("dedisp_select_implementation", C.c_int, [C.c_int]),
    ("dedisp_create_plan", C.c_int, [C.POINTER(C.c_void_p), C.c_ulong, C.c_float, C.c_float, C.c_float]),
    ("dedisp_destroy_plan", None, [C.c_void_p]),
    ("dedisp_set_device", C.c_int, [C.c_int]),
    ("dedisp_set_gulp_size", C.c_int, [C.c_void_p, C.c_ulong]),
    ("dedisp_get_gulp_size", C.c_ulong, [C.c_void_p]),
    ("dedisp_set_killmask", C.c_int, [C.c_void_p, C.c_void_p]),
    ("dedisp_set_dm_list", C.c_int, [C.c_void_p, C.c_void_p, C.c_ulong]),
    ("dedisp_generate_dm_list", C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float]),
    ("dedisp_generate_dm_list_guru", C.c_void_p,
     [C.c_float, C.c_float, C.c_double, C.c_double, C.c_double, C.c_double, C.c_ulong, C.c_double,
      C.POINTER(C.c_ulong)]),
    ("dedisp_free_dm_list_guru", None, [C.c_void_p]),
    ("dedisp_get_max_delay", C.c_ulong, [C.c_void_p]),
    ("dedisp_get_dm_delay", C.c_ulong, [C.c_void_p, C.c_int]),
    ("dedisp_get_channel_count", C.c_ulong, [C.c_void_p]),
    ("dedisp_get_dm_count", C.c_ulong, [C.c_void_p]),
    ("dedisp_get_dm_list", C.c_void_p, [C.c_void_p]),
    ("dedisp_get_killmask", C.c_void_p, [C.c_void_p]),
    ("dedisp_get_dt", C.c_float, [C.c_void_p]),
    ("dedisp_get_f0", C.c_float, [C.c_void_p]),
    ("dedisp_get_df", C.c_float, [C.c_void_p]),
    ("dedisp_get_error_string", C.c_char_p, [C.c_int]),
    ("dedisp_execute", C.c_int, [C.c_void_p, C.c_ulong, C.c_void_p, C.c_ulong, C.c_void_p, C.c_ulong, C.c_uint]),
    ("dedisp_execute_adv", C.c_int,
     [C.c_void_p, C.c_ulong, C.c_void_p, C.c_ulong, C.c_ulong, C.c_void_p, C.c_ulong, C.c_ulong, C.c_uint]),
    ("dedisp_execute_guru", C.c_int,
     [C.c_void_p, C.c_ulong, C.c_void_p, C.c_ulong, C.c_ulong, C.c_void_p, C.c_ulong, C.c_ulong,
      C.c_ulong, C.c_ulong, C.c_uint]),
    ("dedisp_sync", C.c_int, []),
    ("dedisp_enable_adaptive_dt", C.c_int, [C.c_void_p, C.c_float, C.c_float]),
    ("dedisp_disable_adaptive_dt", C.c_int, [C.c_void_p]),
    ("dedisp_using_adaptive_dt", C.c_int, [C.c_void_p]),
    ("dedisp_get_dt_factors", C.c_void_p, [C.c_void_p]),
    ("dedisp_amd_set_profiling", C.c_int, [C.c_void_p, C.c_int]),
    ("dedisp_amd_get_profile", C.c_int, [C.c_void_p, C.POINTER(Profile)]),
    ("dedisp_amd_fdd_set_output", C.c_int, [C.c_void_p, C.c_int]),
    ("dedisp_amd_fdd_set_precision", C.c_int, [C.c_void_p, C.c_int]),
    ("dedisp_amd_fdd_set_segmented", C.c_int, [C.c_void_p, C.c_int, C.c_ulong]),
    ("dedisp_amd_fdd_output_length", C.c_ulong, [C.c_void_p, C.c_ulong]),
    ("dedisp_amd_get_implementation", C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    ("dedisp_amd_generate_delay_table", C.c_int, [C.c_ulong, C.c_float, C.c_float, C.c_float, C.c_void_p]),
    ("dedisp_amd_version", C.c_char_p, []),
    ("dedisp_amd_stream_bytes", C.c_ulong, [C.c_void_p, C.c_ulong, C.c_ulong, C.c_ulong, C.c_ulong]),
    ("dedisp_amd_stream_create", C.c_int, [C.POINTER(C.c_void_p), C.c_void_p, C.c_ulong, C.c_ulong,
                                           C.c_ulong, C.c_ulong]),
    ("dedisp_amd_stream_push", C.c_int, [C.c_void_p, C.c_ulong, C.c_void_p, C.c_void_p, C.c_uint]),
    ("dedisp_amd_stream_push_batch", C.c_int, [C.c_void_p, C.c_ulong, C.POINTER(C.c_ulong),
                                               C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_uint]),
    ("dedisp_amd_stream_reset", C.c_int, [C.c_void_p]),
    ("dedisp_amd_stream_get_stats", C.c_int, [C.c_void_p, C.POINTER(StreamStats)]),
    ("dedisp_amd_stream_history", C.c_ulong, [C.c_void_p]),
    ("dedisp_amd_stream_destroy", None, [C.c_void_p]),
    ("dedisp_amd_fil_read_header", C.c_int, [C.c_char_p, C.POINTER(FilHeader)]),
    ("dedisp_amd_fil_read_data", C.c_int, [C.c_char_p, C.POINTER(FilHeader), C.c_void_p, C.c_ulong]),
    ("dedisp_amd_fil_decimate", C.c_ulong, [C.c_void_p, C.c_ulong, C.c_ulong, C.c_ulong]),
    ("dedisp_amd_write_dat", C.c_int, [C.c_char_p, C.c_float, C.c_void_p, C.c_ulong]),
    ("dedisp_amd_write_inf", C.c_int, [C.c_char_p, C.POINTER(FilHeader), C.c_float, C.c_ulong]),
]
EXPORTED = [s[0] for s in _SIG]


def lib() -> C.CDLL:
    """Loads libdedisp.so (raises if it has not been built).

    PyTorch wheels carry their own copy of the HIP runtime under the same
    soname; when torch is installed it is imported first, so that this library
    binds to the runtime already in the process instead of loading a second
    one (two HIP runtimes in one process cannot share the device)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m dedisp_tests_amd.build`")
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        l = C.CDLL(str(LIB_PATH))
        for name, res, args in _SIG:
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def error_string(code: int) -> str:
    return lib().dedisp_get_error_string(code).decode()


def _check(code: int, where: str) -> None:
    if code != 0:
        raise DedispError(code, where)


def _ptr(a) -> C.c_void_p:
    if a is None:
        return C.c_void_p(0)
    if isinstance(a, int):
        return C.c_void_p(a)
    if hasattr(a, "data_ptr"):  # torch tensor (device memory)
        return C.c_void_p(a.data_ptr())
    return C.c_void_p(a.ctypes.data)


def select_implementation(impl: int) -> None:
    _check(lib().dedisp_select_implementation(impl), "dedisp_select_implementation")


def set_device(idx: int) -> None:
    _check(lib().dedisp_set_device(idx), "dedisp_set_device")


def generate_delay_table(nchans: int, dt: float, f0: float, df: float) -> np.ndarray:
    out = np.empty(nchans, dtype=np.float32)
    _check(lib().dedisp_amd_generate_delay_table(nchans, dt, f0, df, _ptr(out)),
           "dedisp_amd_generate_delay_table")
    return out


def generate_dm_list_guru(dm_start, dm_end, dt, ti, f0, df, nchans, tol) -> np.ndarray:
    n = C.c_ulong(0)
    p = lib().dedisp_generate_dm_list_guru(dm_start, dm_end, dt, ti, f0, df, nchans, tol, C.byref(n))
    if not p:
        raise DedispError(DEDISP_UNKNOWN_ERROR, "dedisp_generate_dm_list_guru")
    arr = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_float)), shape=(n.value,)).copy()
    lib().dedisp_free_dm_list_guru(p)
    return arr


def sync() -> None:
    _check(lib().dedisp_sync(), "dedisp_sync")


class Plan:
    """A dedisp plan (TDD, FDD or the legacy DEDISP selector, which runs TDD)."""

    def __init__(self, nchans: int, dt: float, f0: float, df: float, impl: int = TDD, device: int = 0):
        l = lib()
        _check(l.dedisp_set_device(device), "dedisp_set_device")
        _check(l.dedisp_select_implementation(impl), "dedisp_select_implementation")
        h = C.c_void_p(0)
        _check(l.dedisp_create_plan(C.byref(h), nchans, dt, f0, df), "dedisp_create_plan")
        self._h = h
        self.impl = impl
        self._streams = weakref.WeakSet()  # streams over this plan: closed before it

    def close(self) -> None:
        if getattr(self, "_h", None):
            for st in list(getattr(self, "_streams", ())):
                st.close()
            lib().dedisp_destroy_plan(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    # --- tables ---
    def set_dm_list(self, dms) -> None:
        dms = np.ascontiguousarray(dms, dtype=np.float32)
        _check(lib().dedisp_set_dm_list(self._h, _ptr(dms), dms.size), "dedisp_set_dm_list")

    def generate_dm_list(self, dm_start, dm_end, ti, tol) -> None:
        _check(lib().dedisp_generate_dm_list(self._h, dm_start, dm_end, ti, tol), "dedisp_generate_dm_list")

    def set_killmask(self, killmask) -> None:
        if killmask is None:
            _check(lib().dedisp_set_killmask(self._h, None), "dedisp_set_killmask")
            return
        km = np.ascontiguousarray(killmask, dtype=np.int32)
        _check(lib().dedisp_set_killmask(self._h, _ptr(km)), "dedisp_set_killmask")

    @property
    def max_delay(self) -> int:
        return lib().dedisp_get_max_delay(self._h)

    @property
    def dm_count(self) -> int:
        return lib().dedisp_get_dm_count(self._h)

    @property
    def channel_count(self) -> int:
        return lib().dedisp_get_channel_count(self._h)

    def dm_list(self) -> np.ndarray:
        p = lib().dedisp_get_dm_list(self._h)
        if not p:
            return np.empty(0, np.float32)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_float)), shape=(self.dm_count,)).copy()

    def killmask(self) -> np.ndarray:
        p = lib().dedisp_get_killmask(self._h)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_int)), shape=(self.channel_count,)).copy()

    # --- execution ---
    def out_length(self, nsamps: int) -> int:
        if self.impl == FDD:
            return lib().dedisp_amd_fdd_output_length(self._h, nsamps)
        return nsamps - self.max_delay

    def execute(self, data: np.ndarray, in_nbits: int, out_nbits: int = 32, nsamps: int | None = None,
                flags: int = 0) -> np.ndarray:
        """Host-pointer execute. data: uint8 array, time-major rows of nchans*in_nbits/8 bytes."""
        data = np.ascontiguousarray(data)
        row = self.channel_count * in_nbits // 8
        if nsamps is None:
            nsamps = data.nbytes // row
        ndm = self.dm_count
        nout = self.out_length(nsamps)
        dt = {8: np.uint8, 16: np.uint16, 32: np.float32}[out_nbits]
        out = np.zeros((ndm, nout), dtype=dt)
        _check(lib().dedisp_execute(self._h, nsamps, _ptr(data), in_nbits, _ptr(out), out_nbits, flags),
               "dedisp_execute")
        return out

    def execute_raw(self, nsamps, in_ptr, in_nbits, out_ptr, out_nbits, flags=0) -> int:
        return lib().dedisp_execute(self._h, nsamps, _ptr(in_ptr), in_nbits, _ptr(out_ptr), out_nbits, flags)

    def execute_noalloc_raw(self, nsamps, in_ptr, in_nbits, out_ptr, out_nbits, flags=0) -> int:
        """dedisp_amd_execute_noalloc: device pointers, the library's fixed
        scratch instead of device allocations; returns the dedisp_error code."""
        return lib().dedisp_amd_execute_noalloc(self._h, nsamps, _ptr(in_ptr), in_nbits, _ptr(out_ptr),
                                                out_nbits, flags)

    def execute_adv(self, nsamps, in_ptr, in_nbits, in_stride, out_ptr, out_nbits, out_stride, flags=0) -> int:
        return lib().dedisp_execute_adv(self._h, nsamps, _ptr(in_ptr), in_nbits, in_stride, _ptr(out_ptr),
                                        out_nbits, out_stride, flags)

    def execute_guru(self, nsamps, in_ptr, in_nbits, in_stride, out_ptr, out_nbits, out_stride,
                     first_dm, dm_count, flags=0) -> int:
        return lib().dedisp_execute_guru(self._h, nsamps, _ptr(in_ptr), in_nbits, in_stride, _ptr(out_ptr),
                                         out_nbits, out_stride, first_dm, dm_count, flags)

    # --- extensions ---
    def set_profiling(self, on: bool = True) -> None:
        _check(lib().dedisp_amd_set_profiling(self._h, int(on)), "dedisp_amd_set_profiling")

    def profile(self) -> dict:
        p = Profile()
        _check(lib().dedisp_amd_get_profile(self._h, C.byref(p)), "dedisp_amd_get_profile")
        return p.as_dict()

    def fdd_set_output(self, mode: int) -> None:
        _check(lib().dedisp_amd_fdd_set_output(self._h, mode), "dedisp_amd_fdd_set_output")

    def fdd_set_precision(self, prec: int) -> None:
        _check(lib().dedisp_amd_fdd_set_precision(self._h, prec), "dedisp_amd_fdd_set_precision")

    def fdd_set_segmented(self, enable: bool, nfft: int = 0) -> None:
        _check(lib().dedisp_amd_fdd_set_segmented(self._h, int(enable), nfft), "dedisp_amd_fdd_set_segmented")

    def execute_batch_raw(self, nsamps, in_ptrs, in_nbits, out_ptrs, out_nbits, flags=0) -> int:
        """dedisp_amd_execute_batch over len(nsamps) beams (host arrays or device tensors)."""
        n = len(nsamps)
        ns = (C.c_ulong * n)(*[int(x) for x in nsamps])
        ip = (C.c_void_p * n)(*[_ptr(x).value for x in in_ptrs])
        op = (C.c_void_p * n)(*[_ptr(x).value for x in out_ptrs])
        return lib().dedisp_amd_execute_batch(self._h, n, ns, ip, in_nbits, op, out_nbits, flags)

    def execute_batch(self, beams, in_nbits: int, out_nbits: int = 32, flags: int = 0) -> list:
        """Host-pointer batch: beams is a list of uint8 arrays (time-major rows);
        returns one [ndm, out_length] array per beam."""
        beams = [np.ascontiguousarray(b) for b in beams]
        row = self.channel_count * in_nbits // 8
        nsamps = [b.nbytes // row for b in beams]
        dt = {8: np.uint8, 16: np.uint16, 32: np.float32}[out_nbits]
        outs = [np.zeros((self.dm_count, self.out_length(n)), dtype=dt) for n in nsamps]
        _check(self.execute_batch_raw(nsamps, beams, in_nbits, outs, out_nbits, flags), "dedisp_amd_execute_batch")
        return outs

    def autotune(self, nsamps: int, in_nbits: int, out_nbits: int = 32) -> dict:
        t = Tuning()
        _check(lib().dedisp_amd_autotune(self._h, nsamps, in_nbits, out_nbits, C.byref(t)), "dedisp_amd_autotune")
        return t.as_dict()


def pack_samples(samples: np.ndarray, nbits: int) -> np.ndarray:
    """Packs unsigned samples [nsamps, nchans] into the time-major input layout:
    channel c of a row at bit c*nbits, LSB first (reference src/clib/dedisp.h:338-344).
    Returns uint8 [nsamps, nchans*nbits/8]."""
    s = np.ascontiguousarray(samples)
    nsamps, nchans = s.shape
    if nbits == 8:
        return s.astype(np.uint8)
    if nbits == 16:
        return s.astype("<u2").view(np.uint8).reshape(nsamps, nchans * 2)
    if nbits == 32:
        return s.astype("<u4").view(np.uint8).reshape(nsamps, nchans * 4)
    per = 8 // nbits
    assert nchans % per == 0, "nchans must fill whole bytes"
    v = s.astype(np.uint8).reshape(nsamps, nchans // per, per)
    out = np.zeros((nsamps, nchans // per), dtype=np.uint8)
    for k in range(per):
        out |= (v[:, :, k] & ((1 << nbits) - 1)) << (k * nbits)
    return out


class Stream:
    """Streaming front end over a plan (dedisp_amd_stream_*): nbeams beams that
    deliver chunks of chunk_nsamps samples; each push returns chunk_nsamps
    outputs per DM, delayed by the plan's max_delay (see dedisp_amd.h)."""

    def __init__(self, plan: "Plan", nbeams: int, chunk_nsamps: int, in_nbits: int, out_nbits: int = 32):
        h = C.c_void_p(0)
        _check(lib().dedisp_amd_stream_create(C.byref(h), plan.handle, nbeams, chunk_nsamps, in_nbits, out_nbits),
               "dedisp_amd_stream_create")
        self._h = h
        self.plan = plan
        plan._streams.add(self)
        self.nbeams, self.chunk, self.in_nbits, self.out_nbits = nbeams, chunk_nsamps, in_nbits, out_nbits

    @property
    def history(self) -> int:
        return int(lib().dedisp_amd_stream_history(self._h))

    def push_raw(self, beam: int, chunk_ptr, out_ptr, flags: int = 0) -> int:
        return lib().dedisp_amd_stream_push(self._h, beam, _ptr(chunk_ptr), _ptr(out_ptr), flags)

    def push(self, beam: int, chunk: np.ndarray) -> np.ndarray:
        """Host arrays: chunk uint8 [chunk_nsamps, row_bytes] -> [ndm, chunk_nsamps]."""
        chunk = np.ascontiguousarray(chunk)
        dt = {8: np.uint8, 16: np.uint16, 32: np.float32}[self.out_nbits]
        out = np.empty((self.plan.dm_count, self.chunk), dtype=dt)
        _check(self.push_raw(beam, chunk, out), "dedisp_amd_stream_push")
        return out

    def push_batch_raw(self, beams, chunk_ptrs, out_ptrs, flags: int = 0) -> int:
        n = len(beams)
        b = (C.c_ulong * n)(*beams)
        ins = (C.c_void_p * n)(*[_ptr(x) for x in chunk_ptrs])
        outs = (C.c_void_p * n)(*[_ptr(x) for x in out_ptrs])
        return lib().dedisp_amd_stream_push_batch(self._h, n, b, ins, outs, flags)

    def push_batch(self, beams, chunks) -> list:
        """Host arrays: one chunk per beam (distinct beams) in one batched push;
        returns one [ndm, chunk_nsamps] array per beam."""
        chunks = [np.ascontiguousarray(c) for c in chunks]
        dt = {8: np.uint8, 16: np.uint16, 32: np.float32}[self.out_nbits]
        outs = [np.empty((self.plan.dm_count, self.chunk), dtype=dt) for _ in chunks]
        _check(self.push_batch_raw(list(beams), chunks, outs), "dedisp_amd_stream_push_batch")
        return outs

    def reset(self) -> None:
        _check(lib().dedisp_amd_stream_reset(self._h), "dedisp_amd_stream_reset")

    def stats(self) -> dict:
        st = StreamStats()
        _check(lib().dedisp_amd_stream_get_stats(self._h, C.byref(st)), "dedisp_amd_stream_get_stats")
        return st.as_dict()

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().dedisp_amd_stream_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_filterbank(path) -> tuple[FilHeader, np.ndarray]:
    """SIGPROC filterbank header and samples (uint8 [nsamps, nchans*nbits/8])."""
    h = FilHeader()
    p = str(path).encode()
    _check(lib().dedisp_amd_fil_read_header(p, C.byref(h)), "dedisp_amd_fil_read_header")
    row = h.nchans * h.nbits // 8
    data = np.empty((h.nsamps, row), dtype=np.uint8)
    _check(lib().dedisp_amd_fil_read_data(p, C.byref(h), _ptr(data), data.nbytes), "dedisp_amd_fil_read_data")
    return h, data


def unpack_samples(packed: np.ndarray, nbits: int, nchans: int) -> np.ndarray:
    """Inverse of pack_samples: uint8 [nsamps, row_bytes] -> uint32 [nsamps, nchans]."""
    p = np.ascontiguousarray(packed)
    nsamps = p.shape[0]
    if nbits == 8:
        return p[:, :nchans].astype(np.uint32)
    if nbits == 16:
        return p.view("<u2")[:, :nchans].astype(np.uint32)
    if nbits == 32:
        return p.view("<u4")[:, :nchans].astype(np.uint32)
    per = 8 // nbits
    out = np.empty((nsamps, p.shape[1] * per), dtype=np.uint32)
    for k in range(per):
        out[:, k::per] = (p >> (k * nbits)) & ((1 << nbits) - 1)
    return out[:, :nchans]
