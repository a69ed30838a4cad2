"""ctypes binding of the CPU oracle (oracle/fxref.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product.  Every function restates the
reference (M-AKIMOTOO/fx-corr) file:line named in oracle/fxref.h.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libfxref.so")
_lib = None

MODE_F32 = 0
MODE_F64 = 1
GRID_ANT1 = 0
GRID_ANT2 = 1


def build() -> str:
    """Compile the oracle with its Makefile (gcc only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


class DelayCfg(C.Structure):
    """main.rs:161-182 DelayEvalConfig (table borrowed by pointer)."""

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
    """main.rs:151-159 FrameDelayEntry."""

    _fields_ = [
        ("int1", C.c_int64), ("int2", C.c_int64),
        ("frac1", C.c_double), ("frac2", C.c_double),
        ("fr_lo1", C.c_float * 2), ("fr_lo2", C.c_float * 2),
    ]


class SectorDesc(C.Structure):
    _fields_ = [
        ("fft_len", C.c_int), ("bits1", C.c_int), ("bits2", C.c_int),
        ("shuffle1", C.c_int32 * 32), ("shuffle2", C.c_int32 * 32),
        ("levels1", C.POINTER(C.c_double)), ("nlevels1", C.c_int),
        ("levels2", C.POINTER(C.c_double)), ("nlevels2", C.c_int),
        ("lsb1", C.c_int), ("lsb2", C.c_int),
        ("a1s", C.c_int64), ("a1e", C.c_int64), ("a2s", C.c_int64),
        ("out_grid", C.c_int), ("df_hz", C.c_double),
        ("need_acf", C.c_int), ("need_xcf", C.c_int), ("need_phased", C.c_int),
        ("acf_overlap_only", C.c_int), ("write_raw", C.c_int),
        ("w1", C.c_double), ("w2", C.c_double), ("bit_out", C.c_int),
        ("output_lsb", C.c_int), ("cpu_threads", C.c_int),
    ]


def _p(dt):
    return C.POINTER(dt)


def _declare(L):
    L.fxr_parse_shuffle.argtypes = [C.c_char_p, _p(C.c_int32)]
    L.fxr_parse_shuffle.restype = C.c_int
    L.fxr_parse_levels.argtypes = [C.c_int, C.c_char_p, _p(C.c_double), C.c_int]
    L.fxr_parse_levels.restype = C.c_int
    L.fxr_quantization_mean_power.argtypes = [_p(C.c_double), C.c_int]
    L.fxr_quantization_mean_power.restype = C.c_double
    L.fxr_build_decode_plan.argtypes = [C.c_int, _p(C.c_int32), _p(C.c_double), C.c_int,
                                        _p(C.c_void_p)]
    L.fxr_build_decode_plan.restype = C.c_int
    L.fxr_free_decode_plan.argtypes = [C.c_void_p]
    L.fxr_decode_plan_shuffle_kind.argtypes = [C.c_void_p]
    L.fxr_decode_plan_shuffle_kind.restype = C.c_int
    L.fxr_apply_shuffle_word.argtypes = [C.c_uint32, C.c_void_p]
    L.fxr_apply_shuffle_word.restype = C.c_uint32
    L.fxr_decode_block.argtypes = [_p(C.c_uint8), C.c_size_t, C.c_size_t, C.c_void_p,
                                   _p(C.c_float), C.c_size_t, C.c_int]
    L.fxr_decode_block.restype = C.c_int
    L.fxr_shift_zerofill.argtypes = [_p(C.c_float), C.c_size_t, C.c_int64]
    L.fxr_rfft_f32.argtypes = [C.c_int, _p(C.c_float), _p(C.c_float)]
    L.fxr_rfft_f32.restype = C.c_int
    L.fxr_rfft_f64.argtypes = [C.c_int, _p(C.c_double), _p(C.c_double)]
    L.fxr_rfft_f64.restype = C.c_int
    L.fxr_irfft_f32.argtypes = [C.c_int, _p(C.c_float), _p(C.c_float)]
    L.fxr_irfft_f32.restype = C.c_int
    L.fxr_apply_delay_and_rate_range.argtypes = [
        _p(C.c_float), C.c_size_t, C.c_size_t, C.c_double, C.c_double, C.c_double,
        C.c_double, C.c_double, C.c_int, C.c_size_t, C.c_size_t]
    L.fxr_compute_frame_delay.argtypes = [C.c_int64, _p(DelayCfg), _p(FrameDelay)]
    L.fxr_compute_frame_delay_f64.argtypes = [C.c_int64, _p(DelayCfg), _p(FrameDelay),
                                              _p(C.c_double), _p(C.c_double)]
    L.fxr_split_delay.argtypes = [C.c_double, C.c_double, _p(C.c_int64), _p(C.c_double)]
    L.fxr_carrier_phase.argtypes = [C.c_double, C.c_double, _p(C.c_float)]
    L.fxr_xcf_phase_start_step.argtypes = [C.c_double, C.c_double, C.c_double, C.c_size_t,
                                           C.c_size_t, _p(C.c_float), _p(C.c_float)]
    L.fxr_f64_to_i64_sat.argtypes = [C.c_double]
    L.fxr_f64_to_i64_sat.restype = C.c_int64
    L.fxr_band_alignment.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.c_double, _p(C.c_int64)]
    L.fxr_band_alignment.restype = C.c_int
    L.fxr_quantise_frame.argtypes = [_p(C.c_float), C.c_size_t, C.c_int, _p(C.c_double),
                                     C.c_int, _p(C.c_int32), _p(C.c_uint8), C.c_size_t]
    L.fxr_quantise_frame.restype = C.c_long
    L.fxr_sector_run.argtypes = [
        _p(SectorDesc), _p(C.c_uint8), _p(C.c_uint8), C.c_int64, C.c_int64, _p(DelayCfg),
        C.c_int, C.c_int, _p(C.c_double), _p(C.c_double), _p(C.c_double), _p(C.c_double),
        _p(C.c_uint8), _p(C.c_int64)]
    L.fxr_sector_run.restype = C.c_int


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


# ---------------------------------------------------------------- args.rs ---
def parse_shuffle(text: str) -> np.ndarray:
    out = (C.c_int32 * 32)()
    if lib().fxr_parse_shuffle(text.encode(), out) != 0:
        raise ValueError("invalid shuffle map")
    return np.array(out[:], dtype=np.int32)


def parse_levels(bits: int, text: str) -> np.ndarray:
    cap = max(1, 1 << min(bits, 20)) if bits > 0 else 1
    buf = np.zeros(cap, dtype=np.float64)
    n = lib().fxr_parse_levels(bits, text.encode(), _ptr(buf, C.c_double), cap)
    if n < 0:
        raise ValueError("invalid levels")
    return buf[:n].copy()


def mean_power(levels) -> float:
    lv = np.ascontiguousarray(levels, dtype=np.float64)
    p = lib().fxr_quantization_mean_power(_ptr(lv, C.c_double), len(lv))
    if p < 0:
        raise ValueError("invalid mean power")
    return p


IDENTITY_SHUFFLE = np.arange(32, dtype=np.int32)


# ---------------------------------------------------------------- decode ----
class DecodePlan:
    def __init__(self, bits, shuffle_internal, levels):
        self.bits = bits
        self.shuffle = np.ascontiguousarray(shuffle_internal, dtype=np.int32)
        self.levels = np.ascontiguousarray(levels, dtype=np.float64)
        h = C.c_void_p()
        rc = lib().fxr_build_decode_plan(bits, _ptr(self.shuffle, C.c_int32),
                                         _ptr(self.levels, C.c_double), len(self.levels),
                                         C.byref(h))
        if rc != 0:
            raise ValueError("bad decode plan")
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.fxr_free_decode_plan(self.h)
            self.h = None

    @property
    def shuffle_kind(self) -> int:
        return lib().fxr_decode_plan_shuffle_kind(self.h)

    def shuffle_word(self, w: int) -> int:
        return lib().fxr_apply_shuffle_word(w & 0xFFFFFFFF, self.h)

    def decode(self, raw: bytes | np.ndarray, samples: int, lsb: bool):
        """Returns (ok, float32 array)."""
        r = np.frombuffer(bytes(raw), dtype=np.uint8) if not isinstance(raw, np.ndarray) \
            else np.ascontiguousarray(raw, dtype=np.uint8)
        out = np.zeros(samples, dtype=np.float32)
        if len(r) == 0:
            r = np.zeros(1, dtype=np.uint8)
            n = 0
        else:
            n = len(r)
        rc = lib().fxr_decode_block(_ptr(r, C.c_uint8), n, samples, self.h,
                                    _ptr(out, C.c_float), samples, int(lsb))
        return rc == 0, out


def shift_zerofill(x: np.ndarray, shift: int) -> np.ndarray:
    y = np.ascontiguousarray(x, dtype=np.float32).copy()
    lib().fxr_shift_zerofill(_ptr(y, C.c_float), len(y), int(shift))
    return y


# ---------------------------------------------------------------- fft -------
def rfft_f32(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    n = len(x)
    out = np.zeros(n + 2, dtype=np.float32)
    if lib().fxr_rfft_f32(n, _ptr(x, C.c_float), _ptr(out, C.c_float)) != 0:
        raise ValueError("fft length must be a power of two >= 4")
    return out.view(np.complex64)


def rfft_f64(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = len(x)
    out = np.zeros(n + 2, dtype=np.float64)
    if lib().fxr_rfft_f64(n, _ptr(x, C.c_double), _ptr(out, C.c_double)) != 0:
        raise ValueError("fft length must be a power of two >= 4")
    return out.view(np.complex128)


def irfft_f32(spec: np.ndarray, n: int) -> np.ndarray:
    s = np.ascontiguousarray(spec, dtype=np.complex64).view(np.float32)
    out = np.zeros(n, dtype=np.float32)
    if lib().fxr_irfft_f32(n, _ptr(s, C.c_float), _ptr(out, C.c_float)) != 0:
        raise ValueError("bad length")
    return out


def apply_delay_and_rate(spec, fft_len, df, delay_s, rate=0.0, accel=0.0, t=0.0,
                         is_lsb=False, start=0, end=None):
    s = np.ascontiguousarray(spec, dtype=np.complex64).copy()
    if end is None:
        end = len(s)
    lib().fxr_apply_delay_and_rate_range(_ptr(s.view(np.float32), C.c_float), len(s),
                                         fft_len, df, delay_s, rate, accel, t, int(is_lsb),
                                         start, end)
    return s


# ---------------------------------------------------------------- delay -----
def split_delay(delay_s: float, fs: float):
    n = C.c_int64()
    f = C.c_double()
    lib().fxr_split_delay(delay_s, fs, C.byref(n), C.byref(f))
    return n.value, f.value


def carrier_phase(f_hz: float, tau_s: float) -> np.complex64:
    o = (C.c_float * 2)()
    lib().fxr_carrier_phase(f_hz, tau_s, o)
    return np.complex64(complex(o[0], o[1]))


def xcf_phase_start_step(df, frac1, frac2, sb1, sb2):
    a = (C.c_float * 2)()
    b = (C.c_float * 2)()
    lib().fxr_xcf_phase_start_step(df, frac1, frac2, sb1, sb2, a, b)
    return np.complex64(complex(a[0], a[1])), np.complex64(complex(b[0], b[1]))


def make_delay_cfg(table=None, **kw):
    """Build a DelayCfg; keeps the table alive on the returned object."""
    cfg = DelayCfg()
    for k, v in kw.items():
        setattr(cfg, k, v)
    if table is not None:
        t = np.ascontiguousarray(table, dtype=np.float64)
        cfg._table = t
        cfg.geom_table = _ptr(t, C.c_double)
        cfg.geom_table_len = len(t)
    else:
        cfg.geom_table = None
        cfg.geom_table_len = 0
    return cfg


def frame_delay(idx: int, cfg: DelayCfg) -> FrameDelay:
    out = FrameDelay()
    lib().fxr_compute_frame_delay(idx, C.byref(cfg), C.byref(out))
    return out


def frame_delay_f64(idx: int, cfg: DelayCfg):
    """frame_delay plus the fringe phasors in FP64: (FrameDelay, fr1, fr2)."""
    out = FrameDelay()
    a, b = (C.c_double * 2)(), (C.c_double * 2)()
    lib().fxr_compute_frame_delay_f64(idx, C.byref(cfg), C.byref(out), a, b)
    return out, complex(a[0], a[1]), complex(b[0], b[1])


def band_alignment(fft, fs, c1_mhz, c2_mhz, bw1_mhz, bw2_mhz):
    o = (C.c_int64 * 4)()
    if lib().fxr_band_alignment(fft, fs, c1_mhz, c2_mhz, bw1_mhz, bw2_mhz, o) != 0:
        raise ValueError("No band overlap")
    return tuple(o[:])  # shift_bins, a1s, a1e, a2s


def quantise_frame(x, bits, levels, shuffle_out):
    x = np.ascontiguousarray(x, dtype=np.float32)
    lv = np.ascontiguousarray(levels, dtype=np.float64)
    sh = np.ascontiguousarray(shuffle_out, dtype=np.int32)
    cap = ((len(x) * bits + 31) // 32) * 4 + 4
    out = np.zeros(cap, dtype=np.uint8)
    n = lib().fxr_quantise_frame(_ptr(x, C.c_float), len(x), bits, _ptr(lv, C.c_double),
                                 len(lv), _ptr(sh, C.c_int32), _ptr(out, C.c_uint8), cap)
    if n < 0:
        raise ValueError("quantise failed")
    return out[:n].copy()


# ---------------------------------------------------------------- sector ----
def sector_run(desc: dict, raw1: np.ndarray, raw2: np.ndarray, first_frame: int, nf: int,
               cfg: DelayCfg, mode: int = MODE_F32, threads: int = 1):
    """Returns dict(acc_ph, acc11, acc12 (complex128), acc22, raw_out, n_failed)."""
    d = SectorDesc()
    n = desc["fft_len"]
    half = n // 2 + 1
    d.fft_len = n
    d.bits1, d.bits2 = desc["bits1"], desc["bits2"]
    sh1 = np.ascontiguousarray(desc.get("shuffle1", IDENTITY_SHUFFLE), dtype=np.int32)
    sh2 = np.ascontiguousarray(desc.get("shuffle2", IDENTITY_SHUFFLE), dtype=np.int32)
    d.shuffle1[:] = sh1.tolist()
    d.shuffle2[:] = sh2.tolist()
    lv1 = np.ascontiguousarray(desc["levels1"], dtype=np.float64)
    lv2 = np.ascontiguousarray(desc["levels2"], dtype=np.float64)
    d.levels1, d.nlevels1 = _ptr(lv1, C.c_double), len(lv1)
    d.levels2, d.nlevels2 = _ptr(lv2, C.c_double), len(lv2)
    d.lsb1, d.lsb2 = int(desc.get("lsb1", 0)), int(desc.get("lsb2", 0))
    d.a1s, d.a1e, d.a2s = desc["a1s"], desc["a1e"], desc["a2s"]
    d.out_grid = desc.get("out_grid", GRID_ANT1)
    d.df_hz = desc["df_hz"]
    d.need_acf = int(desc.get("need_acf", 1))
    d.need_xcf = int(desc.get("need_xcf", 1))
    d.need_phased = int(desc.get("need_phased", 0))
    d.acf_overlap_only = int(desc.get("acf_overlap_only", 0))
    d.write_raw = int(desc.get("write_raw", 0))
    d.w1, d.w2 = desc.get("w1", 1.0), desc.get("w2", 1.0)
    d.bit_out = desc.get("bit_out", max(d.bits1, d.bits2))
    d.output_lsb = int(desc.get("output_lsb", 0))
    d.cpu_threads = desc.get("cpu_threads", threads)
    raw1 = np.ascontiguousarray(raw1, dtype=np.uint8)
    raw2 = np.ascontiguousarray(raw2, dtype=np.uint8)
    acc_ph = np.zeros(half)
    acc11 = np.zeros(half)
    acc22 = np.zeros(half)
    acc12 = np.zeros(2 * half)
    bpf_o = (n * d.bit_out + 7) // 8
    raw_out = np.zeros(max(1, nf * bpf_o), dtype=np.uint8)
    nfail = C.c_int64()
    rc = lib().fxr_sector_run(C.byref(d), _ptr(raw1, C.c_uint8), _ptr(raw2, C.c_uint8),
                              first_frame, nf, C.byref(cfg), mode, threads,
                              _ptr(acc_ph, C.c_double), _ptr(acc11, C.c_double),
                              _ptr(acc12, C.c_double), _ptr(acc22, C.c_double),
                              _ptr(raw_out, C.c_uint8), C.byref(nfail))
    if rc != 0:
        raise ValueError("sector_run: bad arguments")
    return dict(acc_ph=acc_ph, acc11=acc11, acc22=acc22,
                acc12=acc12.view(np.complex128), raw_out=raw_out[:nf * bpf_o],
                n_failed=nfail.value)


# --------------------------------------------------- whole-scan cross spectrum
def finalize_cross_spectrum(spec, fft_len, sampling_rate_hz, sky_low_mhz, is_lsb=False):
    """xcf.rs:13-104 (plots left out), restated in plain Python over f64.

    Bins 0 and fft_len/2 (when present) are zeroed (xcf.rs:26-30); per bin
    magnitude (hypot) and phase (safe_arg, utils.rs:425); the DC magnitude is
    replaced by the mean of the others halved (xcf.rs:39-42); the phase of
    bins 1.. is unwrapped among themselves (xcf.rs:44-53, utils.rs:412-424);
    the sky frequency of bin i is sky_low + i df (USB) or sky_low + bw - i df
    (LSB) (xcf.rs:55-63); a least-squares line through (f, phase) over bins
    [1, len) or [1, len - 1) for even fft_len gives delay = -slope / 2 pi
    (xcf.rs:65-101)."""
    import math
    s = [complex(v) for v in spec]
    n = len(s)
    s[0] = 0j
    if fft_len // 2 < n:
        s[fft_len // 2] = 0j
    mag = [math.hypot(v.real, v.imag) for v in s]
    ph = [0.0 if (v.real == 0.0 and v.imag == 0.0) else math.atan2(v.imag, v.real) for v in s]
    if n > 1:
        rest = 0.0
        for v in mag[1:]:
            rest += v
        mag[0] = rest / (n - 1) / 2.0
    unw = list(ph)
    offset = 0.0
    for i in range(2, n):
        d = ph[i] - ph[i - 1]
        if d > math.pi:
            offset -= 2.0 * math.pi
        elif d < -math.pi:
            offset += 2.0 * math.pi
        unw[i] = ph[i] + offset
    bw_mhz = sampling_rate_hz / 2.0 / 1e6
    df_hz = sampling_rate_hz / fft_len
    fhz = []
    for i in range(n):
        fm = (sky_low_mhz + bw_mhz) - (i * df_hz) / 1e6 if is_lsb else sky_low_mhz + (i * df_hz) / 1e6
        fhz.append(fm * 1.0e6)
    delay = slope = icpt = None
    if n > 1:
        end = n - 1 if (fft_len % 2 == 0 and n > 1) else n
        if end > 1:
            xs, ys = fhz[1:end], unw[1:end]
            cnt = float(len(xs))
            sx = sy = sxy = sx2 = 0.0
            for x in xs:
                sx += x
            for y in ys:
                sy += y
            for x, y in zip(xs, ys):
                sxy += x * y
            for x in xs:
                sx2 += x * x
            den = cnt * sx2 - sx * sx
            if abs(den) > 1e-9:
                slope = (cnt * sxy - sx * sy) / den
                icpt = (sy - slope * sx) / cnt
                delay = -slope / (2.0 * math.pi)
    return (np.array(s), dict(delay=delay, slope=slope, intercept=icpt, mag=np.array(mag),
                              phase=np.array(ph), unwrapped=np.array(unw), fhz=np.array(fhz)))


# ------------------------------------------------- auto spectra and phased plots
def finalize_auto_spectrum(acc, power_norm):
    """acf.rs:3-23: acc / power_norm (acc when power_norm <= 0), bin 0 zeroed."""
    out = np.array(acc, dtype=np.float64) / power_norm if power_norm > 0 else np.array(acc, float)
    if out.size:
        out[0] = 0.0
    return out


def phased_plot_series(acc_ph, acc11, acc22, fft_len, emitted, obs_mhz, df_hz):
    """main.rs:2727-2768 without drawing: dict(freqs_mhz, phased_power, amp_ph,
    amp11, amp22, acf_shifted).  The inverse C2C of rustfft x 1/N (utils.rs:30-36)
    is numpy's ifft, which carries the 1/N."""
    norm = max(emitted * fft_len, 1.0)
    ph, a11, a22 = (finalize_auto_spectrum(a, norm) for a in (acc_ph, acc11, acc22))
    bins = fft_len // 2 if fft_len % 2 == 0 else ph.size
    full = np.zeros(fft_len, dtype=np.complex128)
    for i in range(bins):
        v = float(np.float32(ph[i]))
        full[i] = v
        if 0 < i < fft_len // 2:
            full[fft_len - i] = v
    acf = np.fft.ifft(full).real.astype(np.float32).astype(np.float64)
    df = df_hz / 1e6
    return dict(freqs_mhz=obs_mhz + np.arange(bins) * df, phased_power=ph[:bins],
                amp_ph=np.sqrt(ph[:bins]), amp11=np.sqrt(a11[:bins]), amp22=np.sqrt(a22[:bins]),
                acf_shifted=np.roll(acf, -(fft_len // 2)))
