"""Pins the oracle's FFT (realfft method, utils.rs:37-66), delay model
(main.rs:110-241), band alignment (main.rs:360-372) and phase helpers
(main.rs:133-149, utils.rs:93-124) against numpy FP64, closed forms and the
reference README's known answers."""
import math

import numpy as np
import pytest

import fxref as R


@pytest.mark.parametrize("n", [4, 64, 256, 1024, 4096, 65536])
def test_rfft_f32_vs_numpy(n):
    x = np.random.default_rng(n).standard_normal(n).astype(np.float32)
    ref = np.fft.rfft(x.astype(np.float64))
    got = R.rfft_f32(x)
    scale = np.sqrt(n) * np.std(x)
    # FP32 rounding only: rms error ~ 1e-7 * sqrt(log2 n) of the bin scale
    assert np.sqrt(np.mean(np.abs(got - ref) ** 2)) / scale < 3e-7 * np.sqrt(np.log2(n))
    assert got[0].imag == 0 and got[-1].imag == 0   # DC / Nyquist purely real
    got64 = R.rfft_f64(x.astype(np.float64))
    assert np.max(np.abs(got64 - ref)) / scale < 1e-13


def test_rfft_closed_forms():
    n = 1024
    imp = np.zeros(n, dtype=np.float32)
    imp[0] = 1
    assert np.allclose(R.rfft_f32(imp), 1.0, atol=1e-6)
    k = 37
    c = np.cos(2 * np.pi * k * np.arange(n) / n).astype(np.float32)
    s = R.rfft_f32(c)
    assert abs(s[k] - n / 2) < 1e-3
    s[k] = 0
    assert np.max(np.abs(s)) < 1e-3
    # a delayed impulse gives a linear phase exp(-2 pi i k d / n)
    d = 5
    imp = np.zeros(n, dtype=np.float32)
    imp[d] = 1
    kk = np.arange(n // 2 + 1)
    assert np.allclose(R.rfft_f32(imp), np.exp(-2j * np.pi * kk * d / n), atol=1e-6)


@pytest.mark.parametrize("n", [64, 1024, 8192])
def test_irfft_roundtrip_and_scale(n):
    x = np.random.default_rng(1).standard_normal(n).astype(np.float32)
    y = R.irfft_f32(R.rfft_f32(x), n)
    assert np.max(np.abs(y - x)) < 1e-5 * np.sqrt(np.log2(n))
    # against numpy irfft (which ignores DC/Nyquist imaginary parts the same way)
    spec = (np.random.default_rng(2).standard_normal(n // 2 + 1)
            + 1j * np.random.default_rng(3).standard_normal(n // 2 + 1)).astype(np.complex64)
    assert np.allclose(R.irfft_f32(spec, n), np.fft.irfft(spec.astype(np.complex128), n),
                       atol=1e-6)


def test_split_delay_readme_kats():
    # README.md:233-235: 12.983 -> (13, -0.017), 12.345 -> (12, +0.345)
    n, f = R.split_delay(12.983, 1.0)
    assert n == 13 and abs(f + 0.017) < 1e-12
    n, f = R.split_delay(12.345, 1.0)
    assert n == 12 and abs(f - 0.345) < 1e-12
    # half away from zero, and saturating / NaN casts (Rust `as i64`)
    assert R.split_delay(2.5, 1.0)[0] == 3 and R.split_delay(-2.5, 1.0)[0] == -3
    assert R.split_delay(float("nan"), 1.0)[0] == 0
    assert R.split_delay(1e300, 1.0)[0] == 2**63 - 1


def test_carrier_phase_wrap():
    # phase -2 pi f tau wrapped to [-pi, pi) before the f32 cast (main.rs:110-120)
    assert R.carrier_phase(0.0, 123.0) == np.complex64(1 + 0j)
    for f, tau in [(6.7e9, 3.3e-7), (8.4e9, -1.2345e-3), (1e6, 0.25)]:
        ph = -2 * math.pi * f * tau
        w = math.fmod(ph + math.pi, 2 * math.pi)
        if w < 0:
            w += 2 * math.pi
        w -= math.pi
        assert -math.pi <= w < math.pi
        z = R.carrier_phase(f, tau)
        assert abs(complex(z) - complex(math.cos(w), math.sin(w))) < 2e-7
        # the wrapped angle equals the true phase mod 2 pi
        assert abs(math.remainder(w - ph, 2 * math.pi)) < 1e-6 * max(1.0, abs(ph)) * 1e-3 + 1e-9


def _cfg(table=None, **kw):
    base = dict(frame_dt=1024 / 64e6, fs=64e6, time_offset_s=0.0, coarse_delay_s=0.0,
                delay_user_samples=0.0, extra_delay_rate_sps=0.0, clock1_delay_s=0.0,
                clock1_rate_sps=0.0, clock2_delay_s=0.0, clock2_rate_sps=0.0, d_seek=0.0,
                net_d1_base=0.0, total_rate1_base=0.0, total_accel1_base=0.0, net_d2_base=0.0,
                total_rate2_base=0.0, total_accel2_base=0.0, lo1_hz=0.0, lo2_hz=0.0)
    base.update(kw)
    return R.make_delay_cfg(table, **base)


def test_delay_table_linear_interpolation():
    # linear table -> tau(t_mid) is the straight line through it
    fs = 64e6
    table = 1e-7 + 2e-9 * np.arange(3)
    cfg = _cfg(table, fs=fs, frame_dt=1024 / fs)
    for idx in [0, 1, 1000, 62499, 70000, 200000]:
        d = R.frame_delay(idx, cfg)
        t = (idx + 0.5) * 1024 / fs
        t_clamped = min(t, 2.0)
        tau = 1e-7 + 2e-9 * t_clamped
        n = round(tau * fs)
        assert d.int1 == n and abs(d.frac1 - (tau - n / fs)) < 1e-18
        assert d.int2 == 0 and d.frac2 == 0.0


def test_delay_c1_expectation():
    # 64 Msps case: 37.3-sample delay, word-aligned seek of 32 -> int1 = 5, frac1 = 0.3 sample
    fs = 64e6
    cfg = _cfg(np.zeros(2), fs=fs, frame_dt=1024 / fs, delay_user_samples=37.3, d_seek=32 / fs)
    d = R.frame_delay(0, cfg)
    assert d.int1 == 5 and abs(d.frac1 * fs - 0.3) < 1e-6


def test_delay_clock_and_rate_terms():
    fs = 1024e6
    cfg = _cfg(np.full(9, 3e-7), fs=fs, frame_dt=1024 / fs, clock2_delay_s=1.2e-6,
               clock2_rate_sps=1e-12, extra_delay_rate_sps=7.5e-12, coarse_delay_s=1e-9,
               d_seek=1500 / fs)
    for idx in [0, 12345, 999999, 7999999]:
        d = R.frame_delay(idx, cfg)
        t = (idx + 0.5) * 1024 / fs
        full = 3e-7 + 1e-9 + 7.5e-12 * t + (1.2e-6 + 1e-12 * t)
        tau = full - 1500 / fs
        n = round(tau * fs)
        assert d.int1 == n
        assert abs(d.frac1 - (tau - n / fs)) < 1e-17


def test_delay_analytic_branch():
    fs = 512e6
    cfg = _cfg(None, fs=fs, frame_dt=65536 / fs, net_d1_base=2e-8, total_rate1_base=1e-9,
               total_accel1_base=2e-12, d_seek=4 / fs, lo1_hz=1e9)
    d = R.frame_delay(10, cfg)
    t = 10.5 * 65536 / fs
    tau = 2e-8 + 1e-9 * t + 0.5 * 2e-12 * t * t
    assert d.int1 == round(tau * fs)
    assert complex(*d.fr_lo1) == pytest.approx(complex(R.carrier_phase(1e9, tau + 4 / fs)))


@pytest.mark.parametrize("case", [
    # (fft, fs, c1, c2, bw) -> (shift, a1s, a1e, a2s)
    ((1024, 64e6, 132.0, 132.0, 32.0), (0, 0, 512, 0)),
    ((1024, 1024e6, 8512.0, 8512.0, 512.0), (0, 0, 512, 0)),
    ((65536, 512e6, 100.0 + 128, 116.0 + 128, 256.0), (-2048, 2048, 32768, 0)),
    ((65536, 512e6, 116.0 + 128, 100.0 + 128, 256.0), (2048, 0, 30720, 2048)),
])
def test_band_alignment_closed_forms(case):
    args, exp = case
    fft, fs, c1, c2, bw = args
    assert R.band_alignment(fft, fs, c1, c2, bw, bw) == exp


def test_band_alignment_no_overlap():
    with pytest.raises(ValueError):
        R.band_alignment(1024, 64e6, 100.0, 200.0, 32.0, 32.0)


def test_xcf_phase_start_step_closed_form():
    df, f1, f2, s1, s2 = 1e6, 2.3e-10, -0.7e-10, 3, 5
    start, step = R.xcf_phase_start_step(df, f1, f2, s1, s2)
    a = -2 * math.pi * df
    th0 = np.float32(a * f1 * s1 - a * f2 * s2)
    th = np.float32(a * f1 - a * f2)
    assert abs(complex(start) - complex(math.cos(th0), math.sin(th0))) < 1e-7
    assert abs(complex(step) - complex(math.cos(th), math.sin(th))) < 1e-7


def test_delay_rotation_is_linear_phase():
    # apply_delay_and_rate_regular_bins: bin k rotated by exp(-i 2 pi k df tau)
    n = 1024
    df = 64e6 / n
    tau = 0.3 / 64e6
    spec = np.ones(n // 2 + 1, dtype=np.complex64)
    out = R.apply_delay_and_rate(spec, n, df, tau)
    k = np.arange(n // 2 + 1)
    exp = np.exp(-2j * np.pi * k * df * tau)
    exp[-1] = exp[-1].real          # Nyquist imag zeroed
    assert np.max(np.abs(out - exp)) < 5e-5   # f32 recurrence drift over 512 bins
    lsb = R.apply_delay_and_rate(spec, n, df, tau, is_lsb=True)
    assert np.max(np.abs(lsb[:-1] - np.conj(exp[:-1]))) < 5e-5
