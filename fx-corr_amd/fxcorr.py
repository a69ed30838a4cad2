"""ctypes binding of libfxcorr.so (the C ABI declared in include/fxcorr.h).

The Python side only marshals arguments; every computation runs in the HIP
library.  There is no CPU fallback: if the library or a GPU is missing the
calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfxcorr.so")
_lib = None

OK = 0
GRID_ANT1, GRID_ANT2 = 0, 1
PRODUCT_ACF, PRODUCT_XCF, PRODUCT_PHASED, PRODUCT_RAW_OUT, PRODUCT_ACF_OVERLAP_ONLY = 1, 2, 4, 8, 16
PRECISION_PLAIN_FP32, PRECISION_CHOSEN = 0, 1
N_PRODUCTS = 5  # [ph | 11 | 22 | 12re | 12im]

EXPORTED = [
    "fxc_version", "fxc_strerror", "fxc_last_error", "fxc_device_count", "fxc_device_mem_info",
    "fxc_dev_alloc",
    "fxc_dev_free", "fxc_memcpy_h2d", "fxc_memcpy_d2h", "fxc_memset_d", "fxc_memcpy_d2d",
    "fxc_stream_create",
    "fxc_stream_destroy", "fxc_stream_sync", "fxc_event_create", "fxc_event_destroy",
    "fxc_event_record", "fxc_event_elapsed_ms", "fxc_stream_wait_event", "fxc_parse_shuffle",
    "fxc_parse_levels",
    "fxc_quantization_mean_power", "fxc_compute_band_alignment", "fxc_sector_partition",
    "fxc_split_delay", "fxc_plan_create", "fxc_plan_destroy", "fxc_plan_get_info",
    "fxc_sector_run", "fxc_sector_run_device", "fxc_sector_run_device_noalloc",
    "fxc_plan_set_timing", "fxc_plan_get_timing",
    "fxc_decode_frames", "fxc_rfft_frames", "fxc_irfft_frames", "fxc_delay_eval",
    "fxc_quantise_frames", "fxc_synth_station_device", "fxc_synth_spectral_pair_device", "fxc_reader_open", "fxc_reader_read",
    "fxc_reader_close", "fxc_word_aligned_start", "fxc_normalize_sector",
    "fxc_cor_build_header", "fxc_cor_build_sector_header", "fxc_cor_create",
    "fxc_cor_write_sector", "fxc_cor_finalize",
]


class FxcError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"fxcorr status {status}: {msg}")
        self.status = status


class StationDesc(C.Structure):
    _fields_ = [("bits", C.c_int32), ("lsb", C.c_int32), ("shuffle_internal", C.c_int32 * 32),
                ("levels", C.POINTER(C.c_double)), ("nlevels", C.c_int32)]


class PlanDesc(C.Structure):
    _fields_ = [("fft_len", C.c_int32), ("st", StationDesc * 2), ("a1s", C.c_int64),
                ("a1e", C.c_int64), ("a2s", C.c_int64), ("out_grid", C.c_int32),
                ("df_hz", C.c_double), ("products", C.c_uint32), ("w1", C.c_double),
                ("w2", C.c_double), ("bit_out", C.c_int32), ("output_lsb", C.c_int32),
                ("precision", C.c_int32)]


class DelayCfg(C.Structure):
    _fields_ = [
        ("frame_dt", C.c_double), ("fs", C.c_double), ("time_offset_s", C.c_double),
        ("geom_table", C.POINTER(C.c_double)), ("geom_table_len", C.c_int64),
        ("coarse_delay_s", C.c_double), ("delay_user_samples", C.c_double),
        ("extra_delay_rate_sps", C.c_double),
        ("clock1_delay_s", C.c_double), ("clock1_rate_sps", C.c_double),
        ("clock2_delay_s", C.c_double), ("clock2_rate_sps", C.c_double),
        ("d_seek", C.c_double),
        ("net_d1_base", C.c_double), ("total_rate1_base", C.c_double),
        ("total_accel1_base", C.c_double),
        ("net_d2_base", C.c_double), ("total_rate2_base", C.c_double),
        ("total_accel2_base", C.c_double),
        ("lo1_hz", C.c_double), ("lo2_hz", C.c_double),
    ]


class FrameDelay(C.Structure):
    _fields_ = [("int1", C.c_int64), ("int2", C.c_int64), ("frac1", C.c_double),
                ("frac2", C.c_double), ("fr_lo1", C.c_float * 2), ("fr_lo2", C.c_float * 2)]


class BandAlignment(C.Structure):
    _fields_ = [("shift_bins", C.c_int64), ("a1s", C.c_int64), ("a1e", C.c_int64),
                ("a2s", C.c_int64)]


class PlanInfo(C.Structure):
    _fields_ = [("fft_len", C.c_int32), ("half", C.c_int32), ("bpf1", C.c_int32),
                ("bpf2", C.c_int32), ("bpf_out", C.c_int32), ("decode_ok1", C.c_int32),
                ("decode_ok2", C.c_int32), ("fft_form", C.c_int32),
                ("threads_per_frame", C.c_int32), ("device_id", C.c_int32),
                ("variant", C.c_int32), ("grid_divisor", C.c_int32), ("fft_mfma", C.c_int32)]


class CorStation(C.Structure):
    _fields_ = [("name", C.c_char_p), ("code", C.c_uint8), ("ecef_m", C.c_double * 3)]


class CorHeader(C.Structure):
    _fields_ = [("sampling_speed_hz", C.c_int32), ("observing_frequency_hz", C.c_double),
                ("fft_point", C.c_int32), ("number_of_sector_hint", C.c_int32),
                ("clock_reference_unix_sec", C.c_int64), ("source_name", C.c_char_p),
                ("source_ra_rad", C.c_double), ("source_dec_rad", C.c_double)]


class CorClockModel(C.Structure):
    _fields_ = [("sec", C.c_uint32), ("nsec", C.c_uint32), ("delay", C.c_double),
                ("rate", C.c_double), ("acel", C.c_double), ("jerk", C.c_double),
                ("snap", C.c_double)]


class CorSectorModel(C.Structure):
    _fields_ = [("station1", CorClockModel), ("station2", CorClockModel),
                ("amp", C.c_float * 2), ("phs", C.c_int16 * 2)]


def _P(t):
    return C.POINTER(t)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with `python fx-corr_amd/build.py`")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, sz, d = C.c_void_p, C.c_int32, C.c_int64, C.c_size_t, C.c_double
        sig = {
            "fxc_version": ([], C.c_char_p), "fxc_strerror": ([i32], C.c_char_p),
            "fxc_last_error": ([], C.c_char_p), "fxc_device_count": ([_P(i32)], i32),
            "fxc_device_mem_info": ([_P(C.c_uint64), _P(C.c_uint64)], i32),
            "fxc_dev_alloc": ([sz, _P(vp)], i32), "fxc_dev_free": ([vp], i32),
            "fxc_memcpy_h2d": ([vp, vp, sz, vp], i32), "fxc_memcpy_d2h": ([vp, vp, sz, vp], i32),
            "fxc_memset_d": ([vp, C.c_int, sz, vp], i32),
            "fxc_memcpy_d2d": ([vp, vp, sz, vp], i32),
            "fxc_stream_create": ([_P(vp)], i32), "fxc_stream_destroy": ([vp], i32),
            "fxc_stream_sync": ([vp], i32), "fxc_event_create": ([_P(vp)], i32),
            "fxc_event_destroy": ([vp], i32), "fxc_event_record": ([vp, vp], i32),
            "fxc_event_elapsed_ms": ([vp, vp, _P(C.c_float)], i32),
            "fxc_stream_wait_event": ([vp, vp], i32),
            "fxc_parse_shuffle": ([C.c_char_p, _P(i32)], i32),
            "fxc_parse_levels": ([i32, C.c_char_p, _P(d), i32, _P(i32)], i32),
            "fxc_quantization_mean_power": ([_P(d), i32, _P(d)], i32),
            "fxc_compute_band_alignment": ([i32, d, d, d, d, d, _P(BandAlignment)], i32),
            "fxc_sector_partition": ([i64, d, _P(i64), i64, _P(i64)], i32),
            "fxc_split_delay": ([d, d, _P(i64), _P(d)], i32),
            "fxc_plan_create": ([_P(PlanDesc), _P(vp)], i32),
            "fxc_plan_destroy": ([vp], i32),
            "fxc_plan_get_info": ([vp, _P(PlanInfo)], i32),
            "fxc_sector_run": ([vp, vp, sz, vp, sz, i64, i64, _P(DelayCfg), vp, vp, vp, vp, vp,
                                _P(i64)], i32),
            "fxc_sector_run_device": ([vp, vp, vp, i64, i64, _P(DelayCfg), vp, vp, vp], i32),
            "fxc_sector_run_device_noalloc": ([vp, vp, vp, i64, i64, _P(DelayCfg), vp, vp], i32),
            "fxc_plan_set_timing": ([vp, i32], i32),
            "fxc_plan_get_timing": ([vp, _P(d), _P(i64)], i32),
            "fxc_decode_frames": ([vp, i32, vp, sz, i64, vp, vp, _P(i64)], i32),
            "fxc_rfft_frames": ([vp, vp, i64, vp], i32),
            "fxc_irfft_frames": ([vp, vp, i64, vp], i32),
            "fxc_delay_eval": ([_P(DelayCfg), i64, i64, _P(FrameDelay)], i32),
            "fxc_quantise_frames": ([vp, vp, i64, vp], i32),
            "fxc_synth_station_device": ([vp, i64, i32, _P(i32), i32, C.c_uint64, i32, d, i64,
                                          vp], i32),
            "fxc_reader_open": ([C.c_char_p, C.c_uint64, i32, _P(vp)], i32),
            "fxc_reader_read": ([vp, vp, sz, _P(sz)], i32),
            "fxc_reader_close": ([vp], i32),
            "fxc_word_aligned_start": ([C.c_uint64, i32, _P(C.c_uint64), _P(C.c_uint64),
                                        _P(i64)], i32),
            "fxc_normalize_sector": ([vp, i32, i64, d, d, d, d, vp, vp, vp, vp], i32),
            "fxc_cor_build_header": ([_P(CorHeader), _P(CorStation), _P(CorStation), vp], i32),
            "fxc_cor_build_sector_header": ([i64, C.c_float, _P(CorSectorModel), vp], i32),
            "fxc_cor_create": ([C.c_char_p, _P(CorHeader), _P(CorStation), _P(CorStation),
                                _P(vp)], i32),
            "fxc_cor_write_sector": ([vp, i64, C.c_float, vp, i64, _P(CorSectorModel)], i32),
            "fxc_cor_finalize": ([vp], i32),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(status: int):
    if status != OK:
        raise FxcError(status, lib().fxc_last_error().decode(errors="replace"))


def _vp(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


# ------------------------------------------------------------ host helpers
def parse_shuffle(text: str) -> np.ndarray:
    out = (C.c_int32 * 32)()
    check(lib().fxc_parse_shuffle(text.encode(), out))
    return np.array(out[:], dtype=np.int32)


def parse_levels(bits: int, text: str) -> np.ndarray:
    cap = 1 << min(max(bits, 0), 20)
    buf = (C.c_double * cap)()
    n = C.c_int32()
    check(lib().fxc_parse_levels(bits, text.encode(), buf, cap, C.byref(n)))
    return np.array(buf[:n.value])


def mean_power(levels) -> float:
    lv = np.ascontiguousarray(levels, dtype=np.float64)
    p = C.c_double()
    check(lib().fxc_quantization_mean_power(lv.ctypes.data_as(_P(C.c_double)), len(lv),
                                            C.byref(p)))
    return p.value


def band_alignment(fft, fs, c1_mhz, c2_mhz, bw1_mhz, bw2_mhz):
    o = BandAlignment()
    check(lib().fxc_compute_band_alignment(fft, fs, c1_mhz, c2_mhz, bw1_mhz, bw2_mhz,
                                           C.byref(o)))
    return (o.shift_bins, o.a1s, o.a1e, o.a2s)


def sector_partition(total_frames: int, frame_sec: float):
    n = C.c_int64()
    check(lib().fxc_sector_partition(total_frames, frame_sec, None, 0, C.byref(n)))
    buf = (C.c_int64 * max(1, n.value))()
    check(lib().fxc_sector_partition(total_frames, frame_sec, buf, n.value, C.byref(n)))
    return list(buf[:n.value])


def split_delay(delay_s: float, fs: float):
    n = C.c_int64()
    f = C.c_double()
    check(lib().fxc_split_delay(delay_s, fs, C.byref(n), C.byref(f)))
    return n.value, f.value


def device_count() -> int:
    n = C.c_int32()
    st = lib().fxc_device_count(C.byref(n))
    return n.value if st == OK else 0


def device_mem_info():
    """(free, total) bytes of the current device."""
    f, t = C.c_uint64(), C.c_uint64()
    check(lib().fxc_device_mem_info(C.byref(f), C.byref(t)))
    return f.value, t.value


def make_delay_cfg(table=None, **kw) -> DelayCfg:
    cfg = DelayCfg()
    for k, v in kw.items():
        setattr(cfg, k, v)
    if table is not None:
        t = np.ascontiguousarray(table, dtype=np.float64)
        cfg._table = t
        cfg.geom_table = t.ctypes.data_as(_P(C.c_double))
        cfg.geom_table_len = len(t)
    else:
        cfg.geom_table = None
        cfg.geom_table_len = 0
    return cfg


def delay_eval(cfg: DelayCfg, first: int, n: int):
    out = (FrameDelay * max(1, n))()
    check(lib().fxc_delay_eval(C.byref(cfg), first, n, out))
    return out[:n]


IDENTITY_SHUFFLE = np.arange(32, dtype=np.int32)


# --------------------------------------------------------------------- plan
def _plan_desc(fft_len, bits=(2, 2), levels=None, shuffles=None, lsb=(False, False), band=None,
               df_hz=1.0, products=PRODUCT_ACF | PRODUCT_XCF, out_grid=GRID_ANT1, w=(1.0, 1.0),
               bit_out=None, output_lsb=False, precision=PRECISION_PLAIN_FP32, keep=None):
    """fxc_plan_desc from keyword arguments; level arrays are appended to `keep`."""
    d = PlanDesc()
    d.fft_len = fft_len
    for i in range(2):
        lv = np.ascontiguousarray(levels[i], dtype=np.float64)
        if keep is not None:
            keep.append(lv)
        sh = IDENTITY_SHUFFLE if shuffles is None or shuffles[i] is None else shuffles[i]
        d.st[i].bits = bits[i]
        d.st[i].lsb = int(bool(lsb[i]))
        d.st[i].shuffle_internal[:] = [int(x) for x in sh]
        d.st[i].levels = lv.ctypes.data_as(_P(C.c_double))
        d.st[i].nlevels = len(lv)
    if band is None:
        band = (0, 0, fft_len // 2, 0)
    d.a1s, d.a1e, d.a2s = band[1], band[2], band[3]
    d.out_grid = out_grid
    d.df_hz = df_hz
    d.products = products
    d.w1, d.w2 = w
    d.bit_out = bit_out if bit_out is not None else max(bits)
    d.output_lsb = int(bool(output_lsb))
    d.precision = precision
    return d


class Plan:
    """One fxc_plan (per-sector constants + device tables)."""

    def __init__(self, fft_len, **kw):
        self._keep = []
        d = _plan_desc(fft_len, keep=self._keep, **kw)
        self.desc = d
        self.fft_len = fft_len
        self.half = fft_len // 2 + 1
        h = C.c_void_p()
        check(lib().fxc_plan_create(C.byref(d), C.byref(h)))
        self.h = h

    @property
    def info(self) -> PlanInfo:
        """fxc_plan_info, read at each access (tuning may change it)."""
        info = PlanInfo()
        check(lib().fxc_plan_get_info(self.h, C.byref(info)))
        return info

    def close(self):
        if getattr(self, "h", None):
            lib().fxc_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- per-sector hot path on host buffers
    def sector_run(self, raw1, raw2, first_frame, nf, cfg: DelayCfg):
        raw1 = np.ascontiguousarray(raw1, dtype=np.uint8)
        raw2 = np.ascontiguousarray(raw2, dtype=np.uint8)
        h = self.half
        acc_ph, acc11, acc22 = np.zeros(h), np.zeros(h), np.zeros(h)
        acc12 = np.zeros(2 * h)
        raw_out = np.zeros(max(1, nf * self.info.bpf_out), dtype=np.uint8)
        nfail = C.c_int64()
        check(lib().fxc_sector_run(self.h, _vp(raw1), raw1.nbytes, _vp(raw2), raw2.nbytes,
                                   first_frame, nf, C.byref(cfg), _vp(acc_ph), _vp(acc11),
                                   _vp(acc12), _vp(acc22), _vp(raw_out), C.byref(nfail)))
        return dict(acc_ph=acc_ph, acc11=acc11, acc22=acc22, acc12=acc12.view(np.complex128),
                    raw_out=raw_out[:nf * self.info.bpf_out], n_failed=nfail.value)

    def sector_run_device(self, d_raw1, d_raw2, first_frame, nf, cfg, d_acc, d_raw_out=None,
                          stream=None):
        check(lib().fxc_sector_run_device(self.h, C.c_void_p(d_raw1), C.c_void_p(d_raw2),
                                          first_frame, nf, C.byref(cfg), C.c_void_p(d_acc),
                                          C.c_void_p(d_raw_out or 0), C.c_void_p(stream or 0)))

    def sector_run_device_noalloc(self, d_raw1, d_raw2, first_frame, nf, cfg, d_acc, stream=None):
        """fxc_sector_run_device_noalloc: scratch in the library's static device array."""
        check(lib().fxc_sector_run_device_noalloc(self.h, C.c_void_p(d_raw1), C.c_void_p(d_raw2),
                                                  first_frame, nf, C.byref(cfg), C.c_void_p(d_acc),
                                                  C.c_void_p(stream or 0)))

    def set_timing(self, on: bool):
        check(lib().fxc_plan_set_timing(self.h, int(on)))

    def get_timing(self):
        ms = C.c_double()
        n = C.c_int64()
        check(lib().fxc_plan_get_timing(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    # ---- per-function entries
    def decode_frames(self, station, raw, nframes, shifts=None):
        raw = np.ascontiguousarray(raw, dtype=np.uint8)
        out = np.zeros(max(1, nframes * self.fft_len), dtype=np.float32)
        nfail = C.c_int64()
        sh = None
        if shifts is not None:
            sh = np.ascontiguousarray(shifts, dtype=np.int64)
        check(lib().fxc_decode_frames(self.h, station, _vp(raw), raw.nbytes, nframes,
                                      _vp(sh) if sh is not None else None, _vp(out),
                                      C.byref(nfail)))
        return out[:nframes * self.fft_len].reshape(nframes, self.fft_len), nfail.value

    def rfft_frames(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.fft_len)
        out = np.zeros((x.shape[0], self.half), dtype=np.complex64)
        check(lib().fxc_rfft_frames(self.h, _vp(x), x.shape[0], _vp(out)))
        return out

    def irfft_frames(self, spec):
        s = np.ascontiguousarray(spec, dtype=np.complex64).reshape(-1, self.half)
        out = np.zeros((s.shape[0], self.fft_len), dtype=np.float32)
        check(lib().fxc_irfft_frames(self.h, _vp(s), s.shape[0], _vp(out)))
        return out

    def quantise_frames(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.fft_len)
        out = np.zeros((x.shape[0], self.info.bpf_out), dtype=np.uint8)
        check(lib().fxc_quantise_frames(self.h, _vp(x), x.shape[0], _vp(out)))
        return out


# ------------------------------------------------------- device memory helpers
class DeviceBuffer:
    def __init__(self, nbytes: int):
        p = C.c_void_p()
        check(lib().fxc_dev_alloc(nbytes, C.byref(p)))
        self.ptr = p.value
        self.nbytes = nbytes

    def upload(self, arr: np.ndarray, stream=None):
        a = np.ascontiguousarray(arr)
        check(lib().fxc_memcpy_h2d(C.c_void_p(self.ptr), _vp(a), a.nbytes,
                                   C.c_void_p(stream or 0)))

    def download(self, arr: np.ndarray, stream=None):
        check(lib().fxc_memcpy_d2h(_vp(arr), C.c_void_p(self.ptr), arr.nbytes,
                                   C.c_void_p(stream or 0)))
        return arr

    def copy_from(self, src: "DeviceBuffer", nbytes: int, stream=None):
        if nbytes > min(self.nbytes, src.nbytes):
            raise ValueError("copy larger than either buffer")
        check(lib().fxc_memcpy_d2d(C.c_void_p(self.ptr), C.c_void_p(src.ptr), nbytes,
                                   C.c_void_p(stream or 0)))

    def free(self):
        if self.ptr:
            lib().fxc_dev_free(C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        p = C.c_void_p()
        check(lib().fxc_stream_create(C.byref(p)))
        self.ptr = p.value

    def sync(self):
        check(lib().fxc_stream_sync(C.c_void_p(self.ptr)))

    def wait(self, event: "Event"):
        """Later work on this stream waits for the event (fxc_stream_wait_event)."""
        check(lib().fxc_stream_wait_event(C.c_void_p(self.ptr), C.c_void_p(event.ptr)))

    def destroy(self):
        if self.ptr:
            lib().fxc_stream_destroy(C.c_void_p(self.ptr))
            self.ptr = 0


class Event:
    def __init__(self):
        p = C.c_void_p()
        check(lib().fxc_event_create(C.byref(p)))
        self.ptr = p.value

    def record(self, stream: Stream | None):
        check(lib().fxc_event_record(C.c_void_p(self.ptr),
                                     C.c_void_p(stream.ptr if stream else 0)))

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        check(lib().fxc_event_elapsed_ms(C.c_void_p(self.ptr), C.c_void_p(end.ptr), C.byref(ms)))
        return ms.value


class SynthSpectralDesc(C.Structure):
    _fields_ = [("fft_len", C.c_int32), ("bits", C.c_int32 * 2), ("lsb", C.c_int32 * 2),
                ("shuffle_internal", (C.c_int32 * 32) * 2), ("shift_bins", C.c_int64),
                ("rho_continuum", C.c_double), ("rho_line", C.c_double), ("line_bin", C.c_double),
                ("line_fwhm_bins", C.c_double), ("thresh_sigma", C.c_double),
                ("clip_sigma", C.c_double), ("seed", C.c_uint64)]


def synth_spectral_pair_device(d_raw1: int, d_raw2: int, fft_len: int, first_frame: int,
                               nframes: int, bits=(2, 2), lsb=(False, False), shuffles=(None, None),
                               shift_bins=0, rho_continuum=1e-3, rho_line=0.0, line_bin=0.0,
                               line_fwhm_bins=3.0, thresh_sigma=0.98, clip_sigma=4.0, seed=0,
                               stream=None):
    """fxc_synth_spectral_pair_device: seeded pair with a coherence profile
    (continuum + Gaussian line) and a band shift, frame by frame."""
    d = SynthSpectralDesc()
    d.fft_len = fft_len
    for i in range(2):
        d.bits[i] = bits[i]
        d.lsb[i] = int(bool(lsb[i]))
        sh = IDENTITY_SHUFFLE if shuffles[i] is None else shuffles[i]
        d.shuffle_internal[i][:] = [int(x) for x in sh]
    d.shift_bins = shift_bins
    d.rho_continuum, d.rho_line = rho_continuum, rho_line
    d.line_bin, d.line_fwhm_bins = line_bin, line_fwhm_bins
    d.thresh_sigma, d.clip_sigma = thresh_sigma, clip_sigma
    d.seed = seed
    L = lib()
    f = L.fxc_synth_spectral_pair_device
    f.argtypes = [_P(SynthSpectralDesc), C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
    f.restype = C.c_int32
    check(f(C.byref(d), first_frame, nframes, C.c_void_p(d_raw1), C.c_void_p(d_raw2),
            C.c_void_p(stream or 0)))


def synth_station_device(d_out: int, nsamples: int, bits: int, shuffle=None, lsb=False, seed=0,
                         station=0, rho=0.005, delay_samples=0, stream=None):
    sh = (C.c_int32 * 32)(*[int(x) for x in (IDENTITY_SHUFFLE if shuffle is None else shuffle)])
    check(lib().fxc_synth_station_device(C.c_void_p(d_out), nsamples, bits, sh, int(lsb), seed,
                                         station, rho, delay_samples, C.c_void_p(stream or 0)))


# ------------------------------------------------------------- host I/O
class Reader:
    """PackedSampleReader (main.rs:458-516) over the C ABI."""

    def __init__(self, path: str, start_byte: int = 0, bit_offset: int = 0):
        h = C.c_void_p()
        check(lib().fxc_reader_open(os.fsencode(path), start_byte, bit_offset, C.byref(h)))
        self.h = h

    def read(self, n: int):
        out = np.zeros(n, dtype=np.uint8)
        got = C.c_size_t()
        check(lib().fxc_reader_read(self.h, _vp(out), n, C.byref(got)))
        return out, got.value

    def close(self):
        if self.h:
            lib().fxc_reader_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def word_aligned_start(start_samples: int, bits: int):
    """(actual_samples, start_byte, residual_samples), main.rs:1585-1600."""
    a, b, r = C.c_uint64(), C.c_uint64(), C.c_int64()
    check(lib().fxc_word_aligned_start(start_samples, bits, C.byref(a), C.byref(b), C.byref(r)))
    return a.value, b.value, r.value


def normalize_sector(acc, fft_len, nf, p1, p2, w1=1.0, w2=1.0):
    """Normalised .cor spectra (main.rs:2645-2683) from a [5][half] accumulator."""
    a = np.ascontiguousarray(acc, dtype=np.float64).reshape(-1)
    outs = [np.zeros(fft_len, dtype=np.float32) for _ in range(4)]
    check(lib().fxc_normalize_sector(_vp(a), fft_len, nf, p1, p2, w1, w2,
                                     *[_vp(o) for o in outs]))
    return {k: o.view(np.complex64) for k, o in zip(("ph", "11", "22", "12"), outs)}


def cor_station(name: str, code: int, ecef):
    s = CorStation()
    s.name = name.encode()
    s.code = code
    s.ecef_m[:] = [float(v) for v in ecef]
    return s


def cor_header(fs_hz, obs_hz, fft_point, nsector_hint, clock_ref_unix, source, ra, dec):
    h = CorHeader()
    h.sampling_speed_hz = int(fs_hz)
    h.observing_frequency_hz = float(obs_hz)
    h.fft_point = fft_point
    h.number_of_sector_hint = nsector_hint
    h.clock_reference_unix_sec = clock_ref_unix
    h.source_name = source.encode()
    h.source_ra_rad = ra
    h.source_dec_rad = dec
    return h


def cor_build_header(hdr: CorHeader, st1: CorStation, st2: CorStation) -> bytes:
    out = np.zeros(256, dtype=np.uint8)
    check(lib().fxc_cor_build_header(C.byref(hdr), C.byref(st1), C.byref(st2), _vp(out)))
    return out.tobytes()


def cor_build_sector_header(unix_sec: int, integ_s: float, model=None) -> bytes:
    out = np.zeros(128, dtype=np.uint8)
    check(lib().fxc_cor_build_sector_header(unix_sec, integ_s,
                                            C.byref(model) if model is not None else None,
                                            _vp(out)))
    return out.tobytes()


class CorWriter:
    """CorWriter (cor.rs:206-296) over the C ABI."""

    def __init__(self, path: str, hdr: CorHeader, st1: CorStation, st2: CorStation):
        h = C.c_void_p()
        check(lib().fxc_cor_create(os.fsencode(path), C.byref(hdr), C.byref(st1),
                                   C.byref(st2), C.byref(h)))
        self.h = h

    def write_sector(self, unix_sec: int, integ_s: float, spectrum, model=None):
        s = np.ascontiguousarray(spectrum, dtype=np.complex64)
        check(lib().fxc_cor_write_sector(self.h, unix_sec, integ_s, _vp(s), s.size,
                                         C.byref(model) if model is not None else None))

    def finalize(self):
        h, self.h = self.h, None
        if h:
            check(lib().fxc_cor_finalize(h))


# ------------------------------------------------------------------ serving
class ServerCfg(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_active_jobs", C.c_int32),
                ("n_streams", C.c_int32)]


class JobDesc(C.Structure):
    _fields_ = [("plan", PlanDesc), ("delay", DelayCfg), ("first_frame", C.c_int64),
                ("input_on_device", C.c_int32)]


class JobStats(C.Structure):
    _fields_ = [("queue_ms", C.c_double), ("latency_ms", C.c_double), ("frames", C.c_int64),
                ("footprint_bytes", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [("jobs_done", C.c_int64), ("jobs_failed", C.c_int64),
                ("p50_latency_ms", C.c_double), ("p99_latency_ms", C.c_double),
                ("mean_latency_ms", C.c_double), ("max_concurrent_jobs", C.c_int32),
                ("peak_hbm_bytes", C.c_uint64)]


def _server_sigs(L):
    vp, i32, i64, d, u64 = C.c_void_p, C.c_int32, C.c_int64, C.c_double, C.c_uint64
    sig = {
        "fxc_server_create": ([_P(ServerCfg), _P(vp)], i32),
        "fxc_server_destroy": ([vp], i32),
        "fxc_job_submit": ([vp, _P(JobDesc), _P(u64)], i32),
        "fxc_job_push": ([vp, u64, vp, vp, i64], i32),
        "fxc_job_close": ([vp, u64], i32),
        "fxc_job_wait": ([vp, u64, d, _P(i32)], i32),
        "fxc_job_result": ([vp, u64, vp, _P(JobStats)], i32),
        "fxc_job_release": ([vp, u64], i32),
        "fxc_server_get_stats": ([vp, _P(ServerStats)], i32),
        "fxc_plan_device_bytes": ([vp, _P(u64)], i32),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res


EXPORTED += ["fxc_server_create", "fxc_server_destroy", "fxc_job_submit", "fxc_job_push",
             "fxc_job_close", "fxc_job_wait", "fxc_job_result", "fxc_job_release",
             "fxc_server_get_stats", "fxc_plan_device_bytes"]


def _slib():
    L = lib()
    if not getattr(L, "_server_sigs", False):
        _server_sigs(L)
        L._server_sigs = True
    return L


class Server:
    """Continuous-batching job server (fxc_server_*)."""

    def __init__(self, hbm_budget_bytes=200 << 30, max_active_jobs=8, n_streams=4):
        cfg = ServerCfg(hbm_budget_bytes, max_active_jobs, n_streams)
        h = C.c_void_p()
        check(_slib().fxc_server_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self._keep = {}

    def submit(self, plan_kw: dict, cfg: DelayCfg, first_frame=0, input_on_device=True) -> int:
        """plan_kw: the Plan() keyword arguments (fft_len, bits, levels, ...)."""
        jd = JobDesc()
        keep = []
        jd.plan = _plan_desc(keep=keep, **plan_kw)
        jd.delay = cfg
        jd.first_frame = first_frame
        jd.input_on_device = int(bool(input_on_device))
        jid = C.c_uint64()
        check(_slib().fxc_job_submit(self.h, C.byref(jd), C.byref(jid)))
        self._keep[jid.value] = []
        return jid.value

    def push(self, jid: int, raw1, raw2, nframes: int):
        """raw1/raw2: device pointers (int) or host numpy arrays kept alive here."""
        if isinstance(raw1, np.ndarray):
            a1, a2 = np.ascontiguousarray(raw1), np.ascontiguousarray(raw2)
            self._keep[jid] += [a1, a2]
            p1, p2 = _vp(a1), _vp(a2)
        else:
            p1, p2 = C.c_void_p(raw1), C.c_void_p(raw2)
        check(_slib().fxc_job_push(self.h, jid, p1, p2, nframes))

    def close_job(self, jid: int):
        check(_slib().fxc_job_close(self.h, jid))

    def wait(self, jid: int, timeout=-1.0) -> bool:
        done = C.c_int32()
        check(_slib().fxc_job_wait(self.h, jid, timeout, C.byref(done)))
        return bool(done.value)

    def result(self, jid: int, half: int):
        acc = np.zeros(N_PRODUCTS * half)
        st = JobStats()
        check(_slib().fxc_job_result(self.h, jid, _vp(acc), C.byref(st)))
        return acc, st

    def release(self, jid: int):
        check(_slib().fxc_job_release(self.h, jid))
        self._keep.pop(jid, None)

    def stats(self) -> ServerStats:
        st = ServerStats()
        check(_slib().fxc_server_get_stats(self.h, C.byref(st)))
        return st

    def close(self):
        if self.h:
            check(_slib().fxc_server_destroy(self.h))
            self.h = None
            self._keep.clear()


def plan_device_bytes(plan: Plan) -> int:
    b = C.c_uint64()
    check(_slib().fxc_plan_device_bytes(plan.h, C.byref(b)))
    return b.value


# ------------------------------------------------------------------ graphs
EXPORTED += ["fxc_graph_begin", "fxc_graph_end", "fxc_graph_launch", "fxc_graph_destroy",
             "fxc_plan_live_graphs", "fxc_plan_check_device_error"]


def _glib():
    L = lib()
    if not getattr(L, "_graph_sigs", False):
        vp, i32 = C.c_void_p, C.c_int32
        for name, args in (("fxc_graph_begin", [vp]), ("fxc_graph_end", [vp, _P(vp)]),
                           ("fxc_graph_launch", [vp, vp]), ("fxc_graph_destroy", [vp]),
                           ("fxc_plan_live_graphs", [vp, _P(i32)]),
                           ("fxc_plan_check_device_error", [vp, vp])):
            f = getattr(L, name)
            f.argtypes = args
            f.restype = i32
        L._graph_sigs = True
    return L


class Graph:
    """Capture of device-entry calls on one stream, replayable (fxc_graph_*)."""

    def __init__(self, stream: Stream):
        self.stream = stream
        self.h = None
        check(_glib().fxc_graph_begin(stream.ptr))

    def end(self):
        h = C.c_void_p()
        check(_glib().fxc_graph_end(self.stream.ptr, C.byref(h)))
        self.h = h
        return self

    def launch(self, stream=None):
        check(_glib().fxc_graph_launch(self.h, (stream or self.stream).ptr))

    def close(self):
        if self.h:
            _glib().fxc_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def plan_live_graphs(plan: "Plan") -> int:
    n = C.c_int32()
    check(_glib().fxc_plan_live_graphs(plan.h, C.byref(n)))
    return n.value


def plan_check_device_error(plan: "Plan", stream=None):
    """Raises FxcError when a fused kernel's bounded hand-over wait ran out."""
    check(_glib().fxc_plan_check_device_error(plan.h, stream.ptr if stream else None))


# ---------------------------------------------------------------- autotune
class TuneReport(C.Structure):
    _fields_ = [("candidates", C.c_int32), ("variant", C.c_int32), ("grid_divisor", C.c_int32),
                ("split_m1", C.c_int32), ("chunk_frames", C.c_int64), ("best_ms", C.c_double),
                ("default_ms", C.c_double)]


EXPORTED += ["fxc_plan_autotune"]


def plan_autotune(plan: Plan, d_raw1: int, d_raw2: int, nf: int, cfg: DelayCfg, stream=None):
    L = lib()
    f = L.fxc_plan_autotune
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, _P(DelayCfg), C.c_void_p,
                  _P(TuneReport)]
    f.restype = C.c_int32
    rep = TuneReport()
    check(f(plan.h, C.c_void_p(d_raw1), C.c_void_p(d_raw2), nf, C.byref(cfg),
            C.c_void_p(stream or 0), C.byref(rep)))
    return {k: getattr(rep, k) for k, _ in TuneReport._fields_}


EXPORTED += ["fxc_plan_set_tuning"]


def plan_set_tuning(plan: Plan, variant=0, grid_divisor=1, split_m1=0, chunk_frames=0):
    L = lib()
    f = L.fxc_plan_set_tuning
    f.argtypes = [C.c_void_p, _P(TuneReport)]
    f.restype = C.c_int32
    r = TuneReport()
    r.variant, r.grid_divisor, r.split_m1, r.chunk_frames = variant, grid_divisor, split_m1, chunk_frames
    check(f(plan.h, C.byref(r)))


# ------------------------------------------------- whole-scan cross spectrum
class CrossSpectrumResult(C.Structure):
    _fields_ = [("has_delay", C.c_int32), ("delay_seconds", C.c_double),
                ("slope_rad_per_hz", C.c_double), ("intercept_rad", C.c_double)]


EXPORTED += ["fxc_fringe_accumulate", "fxc_fringe_accumulate_device",
             "fxc_finalize_cross_spectrum"]


def _fringe_sigs(L):
    if not getattr(L, "_fringe_sigs", False):
        vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
        L.fxc_fringe_accumulate.argtypes = [vp, vp, C.c_size_t, vp, C.c_size_t, i64, i64,
                                            _P(DelayCfg), vp, _P(i64)]
        L.fxc_fringe_accumulate.restype = i32
        L.fxc_fringe_accumulate_device.argtypes = [vp, vp, vp, i64, i64, _P(DelayCfg), vp, vp]
        L.fxc_fringe_accumulate_device.restype = i32
        L.fxc_finalize_cross_spectrum.argtypes = [vp, i64, i32, C.c_double, C.c_double, i32,
                                                  _P(CrossSpectrumResult), vp, vp, vp, vp]
        L.fxc_finalize_cross_spectrum.restype = i32
        L._fringe_sigs = True
    return L


def fringe_accumulate(plan: Plan, raw1, raw2, first_frame, nf, cfg: DelayCfg, ac: np.ndarray):
    """Adds one chunk's XCF sum / N^2 into ac (complex128, plan.half bins);
    returns the number of frames that failed to decode."""
    L = _fringe_sigs(lib())
    raw1 = np.ascontiguousarray(raw1, dtype=np.uint8)
    raw2 = np.ascontiguousarray(raw2, dtype=np.uint8)
    if ac.dtype != np.complex128 or ac.size != plan.half or not ac.flags.c_contiguous:
        raise ValueError("ac must be a contiguous complex128 array of plan.half bins")
    nfail = C.c_int64()
    check(L.fxc_fringe_accumulate(plan.h, _vp(raw1), raw1.nbytes, _vp(raw2), raw2.nbytes,
                                  first_frame, nf, C.byref(cfg), _vp(ac), C.byref(nfail)))
    return nfail.value


def fringe_accumulate_device(plan: Plan, d_raw1, d_raw2, first_frame, nf, cfg: DelayCfg, d_ac,
                             stream=None):
    L = _fringe_sigs(lib())
    check(L.fxc_fringe_accumulate_device(plan.h, C.c_void_p(d_raw1), C.c_void_p(d_raw2),
                                         first_frame, nf, C.byref(cfg), C.c_void_p(d_ac),
                                         C.c_void_p(stream or 0)))


def finalize_cross_spectrum(spec: np.ndarray, fft_len: int, sampling_rate_hz: float,
                            sky_low_mhz: float, is_lsb: bool = False):
    """xcf.rs:13-104 without plots.  Returns (spectrum with bins 0 and
    fft_len/2 zeroed, dict(delay, slope, intercept, mag, phase, unwrapped,
    fit)); delay is None when the fit is degenerate."""
    L = _fringe_sigs(lib())
    s = np.array(spec, dtype=np.complex128, copy=True)
    n = s.size
    mag, ph, unw, fit = (np.zeros(n) for _ in range(4))
    r = CrossSpectrumResult()
    check(L.fxc_finalize_cross_spectrum(_vp(s), n, fft_len, sampling_rate_hz, sky_low_mhz,
                                        int(bool(is_lsb)), C.byref(r), _vp(mag), _vp(ph),
                                        _vp(unw), _vp(fit)))
    return s, dict(delay=r.delay_seconds if r.has_delay else None, slope=r.slope_rad_per_hz,
                   intercept=r.intercept_rad, mag=mag, phase=ph, unwrapped=unw, fit=fit)


# ------------------------------------------------------ streaming host input
EXPORTED += ["fxc_host_alloc", "fxc_host_free", "fxc_pipeline_create", "fxc_pipeline_submit",
             "fxc_pipeline_wait", "fxc_pipeline_destroy"]


def _pipe_sigs(L):
    if not getattr(L, "_pipe_sigs", False):
        vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
        L.fxc_host_alloc.argtypes = [C.c_size_t, _P(vp)]
        L.fxc_host_alloc.restype = i32
        L.fxc_host_free.argtypes = [vp]
        L.fxc_host_free.restype = i32
        L.fxc_pipeline_create.argtypes = [vp, i64, i32, _P(vp)]
        L.fxc_pipeline_create.restype = i32
        L.fxc_pipeline_submit.argtypes = [vp, vp, vp, i64, i64, _P(DelayCfg), vp, _P(i64)]
        L.fxc_pipeline_submit.restype = i32
        L.fxc_pipeline_wait.argtypes = [vp, i64]
        L.fxc_pipeline_wait.restype = i32
        L.fxc_pipeline_destroy.argtypes = [vp]
        L.fxc_pipeline_destroy.restype = i32
        L._pipe_sigs = True
    return L


class HostBuffer:
    """Page-locked host memory (fxc_host_alloc) viewed as a numpy array."""

    def __init__(self, nbytes: int):
        L = _pipe_sigs(lib())
        p = C.c_void_p()
        check(L.fxc_host_alloc(nbytes, C.byref(p)))
        self.ptr, self.nbytes = p.value, nbytes
        self.array = np.ctypeslib.as_array((C.c_uint8 * nbytes).from_address(self.ptr))

    def view(self, dtype):
        return self.array.view(dtype)

    def free(self):
        if self.ptr:
            _pipe_sigs(lib()).fxc_host_free(C.c_void_p(self.ptr))
            self.ptr = 0
            self.array = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Pipeline:
    """fxc_pipeline: sectors from host memory, copies overlapping kernels."""

    def __init__(self, plan: Plan, max_frames: int, depth: int = 2):
        L = _pipe_sigs(lib())
        h = C.c_void_p()
        check(L.fxc_pipeline_create(plan.h, max_frames, depth, C.byref(h)))
        self.h, self.plan = h, plan

    def submit(self, raw1_ptr: int, raw2_ptr: int, first_frame: int, nf: int, cfg: DelayCfg,
               acc_out_ptr: int) -> int:
        t = C.c_int64()
        check(_pipe_sigs(lib()).fxc_pipeline_submit(self.h, C.c_void_p(raw1_ptr),
                                                    C.c_void_p(raw2_ptr), first_frame, nf,
                                                    C.byref(cfg), C.c_void_p(acc_out_ptr),
                                                    C.byref(t)))
        return t.value

    def wait(self, ticket: int):
        check(_pipe_sigs(lib()).fxc_pipeline_wait(self.h, ticket))

    def close(self):
        if getattr(self, "h", None):
            _pipe_sigs(lib()).fxc_pipeline_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------ station geometry
EXPORTED += ["fxc_parse_ra", "fxc_parse_dec", "fxc_parse_epoch_mjd", "fxc_precess_j2000_to_date",
             "fxc_sidereal_time", "fxc_geometric_delay", "fxc_geometric_delay_table",
             "fxc_geometric_table_points", "fxc_geometric_model_bound"]


class GeometricDelay(C.Structure):
    _fields_ = [("delay1_s", C.c_double), ("delay2_s", C.c_double), ("delay_s", C.c_double),
                ("rate_sps", C.c_double), ("accel_sps2", C.c_double)]


def _geom_sigs(L):
    if not getattr(L, "_geom_sigs", False):
        d, pd, i64, i32 = C.c_double, _P(C.c_double), C.c_int64, C.c_int32
        for name in ("fxc_parse_ra", "fxc_parse_dec", "fxc_parse_epoch_mjd"):
            getattr(L, name).argtypes = [C.c_char_p, pd]
            getattr(L, name).restype = i32
        L.fxc_precess_j2000_to_date.argtypes = [d, d, d, pd, pd]
        L.fxc_precess_j2000_to_date.restype = i32
        L.fxc_sidereal_time.argtypes = [d, pd, pd]
        L.fxc_sidereal_time.restype = i32
        L.fxc_geometric_delay.argtypes = [pd, pd, d, d, d, _P(GeometricDelay)]
        L.fxc_geometric_delay.restype = i32
        L.fxc_geometric_delay_table.argtypes = [pd, pd, d, d, d, d, i64, pd]
        L.fxc_geometric_delay_table.restype = i32
        L.fxc_geometric_table_points.argtypes = [i64, d]
        L.fxc_geometric_table_points.restype = i64
        L.fxc_geometric_model_bound.argtypes = [pd, pd, d, pd, pd]
        L.fxc_geometric_model_bound.restype = i32
        L._geom_sigs = True
    return L


def _parse(name, text):
    out = C.c_double()
    check(getattr(_geom_sigs(lib()), name)(text.encode(), C.byref(out)))
    return out.value


def parse_ra(text: str) -> float:
    """geom.rs:86-111: radians."""
    return _parse("fxc_parse_ra", text)


def parse_dec(text: str) -> float:
    """geom.rs:115-148: radians."""
    return _parse("fxc_parse_dec", text)


def parse_epoch_mjd(text: str) -> float:
    """geom.rs:226-257: MJD."""
    return _parse("fxc_parse_epoch_mjd", text)


def precess_j2000_to_date(ra: float, dec: float, mjd: float):
    """geom.rs:265-277: (ra, dec) of date."""
    r, d = C.c_double(), C.c_double()
    check(_geom_sigs(lib()).fxc_precess_j2000_to_date(ra, dec, mjd, C.byref(r), C.byref(d)))
    return r.value, d.value


def sidereal_time(mjd: float):
    """(GMST, GAST) in radians (geom.rs:279-297)."""
    g, a = C.c_double(), C.c_double()
    check(_geom_sigs(lib()).fxc_sidereal_time(mjd, C.byref(g), C.byref(a)))
    return g.value, a.value


def _ecef(x):
    a = (C.c_double * 3)(*[float(v) for v in x])
    return a


def geometric_delay(ant1_ecef, ant2_ecef, ra: float, dec: float, mjd: float):
    """geom.rs:346-384: (delay1, delay2, delay, rate, accel)."""
    r = GeometricDelay()
    check(_geom_sigs(lib()).fxc_geometric_delay(_ecef(ant1_ecef), _ecef(ant2_ecef), ra, dec, mjd,
                                                C.byref(r)))
    return r.delay1_s, r.delay2_s, r.delay_s, r.rate_sps, r.accel_sps2


def geometric_table_points(total_frames: int, frame_dt: float) -> int:
    return int(_geom_sigs(lib()).fxc_geometric_table_points(total_frames, frame_dt))


def geometric_delay_table(ant1_ecef, ant2_ecef, ra: float, dec: float, mjd: float,
                          time_offset_s: float, n_points: int) -> np.ndarray:
    """main.rs:1628-1642: baseline delay at each whole second of the process."""
    out = np.zeros(n_points)
    check(_geom_sigs(lib()).fxc_geometric_delay_table(
        _ecef(ant1_ecef), _ecef(ant2_ecef), ra, dec, mjd, time_offset_s, n_points,
        out.ctypes.data_as(_P(C.c_double))))
    return out


# ------------------------------------------------ auto spectra and phased plots
EXPORTED += ["fxc_finalize_auto_spectrum", "fxc_phased_plot_series"]


class PhasedPlot(C.Structure):
    _fields_ = [("freqs_mhz", _P(C.c_double)), ("phased_power", _P(C.c_double)),
                ("amp_ph", _P(C.c_double)), ("amp11", _P(C.c_double)), ("amp22", _P(C.c_double)),
                ("acf_shifted", _P(C.c_double)), ("spec_bins", C.c_int64)]


def _auto_sigs(L):
    if not getattr(L, "_auto_sigs", False):
        pd = _P(C.c_double)
        L.fxc_finalize_auto_spectrum.argtypes = [pd, C.c_int64, C.c_double, pd]
        L.fxc_finalize_auto_spectrum.restype = C.c_int32
        L.fxc_phased_plot_series.argtypes = [pd, pd, pd, C.c_int64, C.c_int32, C.c_int64,
                                             C.c_double, C.c_double, _P(PhasedPlot)]
        L.fxc_phased_plot_series.restype = C.c_int32
        L._auto_sigs = True
    return L


def _dp(a):
    return a.ctypes.data_as(_P(C.c_double))


def finalize_auto_spectrum(acc, power_norm: float) -> np.ndarray:
    """acf.rs:3-23."""
    a = np.ascontiguousarray(acc, dtype=np.float64)
    out = np.zeros_like(a)
    check(_auto_sigs(lib()).fxc_finalize_auto_spectrum(_dp(a), a.size, power_norm, _dp(out)))
    return out


def phased_plot_series(acc_ph, acc11, acc22, fft_len: int, emitted_frames: int, obs_mhz: float,
                       df_hz: float) -> dict:
    """main.rs:2727-2768 without drawing (see fxc_phased_plot_series)."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (acc_ph, acc11, acc22)]
    h = fft_len // 2 + 1
    outs = {k: np.zeros(h) for k in ("freqs_mhz", "phased_power", "amp_ph", "amp11", "amp22")}
    acf = np.zeros(fft_len)
    pp = PhasedPlot(*[_dp(outs[k]) for k in ("freqs_mhz", "phased_power", "amp_ph", "amp11",
                                             "amp22")], _dp(acf), 0)
    check(_auto_sigs(lib()).fxc_phased_plot_series(_dp(arrs[0]), _dp(arrs[1]), _dp(arrs[2]),
                                                   arrs[0].size, fft_len, emitted_frames,
                                                   obs_mhz, df_hz, C.byref(pp)))
    b = pp.spec_bins
    res = {k: v[:b] for k, v in outs.items()}
    res["acf_shifted"] = acf
    return res


def geometric_model_bound(ecef1, ecef2, obs_hz: float):
    """(delay_s, phase_rad): how far the 2000B-based geometry may move the
    fringe phase from the reference's 2000A-based one on this baseline."""
    dl, ph = C.c_double(), C.c_double()
    check(_geom_sigs(lib()).fxc_geometric_model_bound(_ecef(ecef1), _ecef(ecef2), obs_hz,
                                                      C.byref(dl), C.byref(ph)))
    return dl.value, ph.value
