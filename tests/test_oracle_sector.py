"""Pins the oracle's per-sector compute (main.rs:2173-2591) with invariants
and an independent numpy re-run of the same algorithm."""
import numpy as np
import pytest

import fxref as R
from _synth import IDENT, LEVELS_2BIT, make_pair_raw


def _zero_cfg(fs, n, **kw):
    base = dict(frame_dt=n / fs, fs=fs, time_offset_s=0.0, coarse_delay_s=0.0,
                delay_user_samples=0.0, extra_delay_rate_sps=0.0, clock1_delay_s=0.0,
                clock1_rate_sps=0.0, clock2_delay_s=0.0, clock2_rate_sps=0.0, d_seek=0.0,
                net_d1_base=0.0, total_rate1_base=0.0, total_accel1_base=0.0, net_d2_base=0.0,
                total_rate2_base=0.0, total_accel2_base=0.0, lo1_hz=0.0, lo2_hz=0.0)
    base.update(kw)
    return R.make_delay_cfg(np.zeros(4), **base)


def _desc(n, fs, **kw):
    d = dict(fft_len=n, bits1=2, bits2=2, levels1=LEVELS_2BIT, levels2=LEVELS_2BIT,
             a1s=0, a1e=n // 2, a2s=0, df_hz=fs / n, need_acf=1, need_xcf=1, cpu_threads=2)
    d.update(kw)
    return d


def test_identical_streams_xcf_equals_acf():
    n, fs, nf = 256, 64e6, 64
    raw, _, _ = make_pair_raw(nf, n, seed=1)
    cfg = _zero_cfg(fs, n)
    for mode in (R.MODE_F32, R.MODE_F64):
        r = R.sector_run(_desc(n, fs), raw, raw, 0, nf, cfg, mode=mode, threads=2)
        assert r["n_failed"] == 0
        ov = slice(0, n // 2)
        assert np.allclose(r["acc12"][ov].real, r["acc11"][ov], rtol=1e-6)
        assert np.max(np.abs(r["acc12"][ov].imag)) <= 1e-6 * np.max(r["acc11"])
        # ACF22 covers only the overlap bins [0, n/2): Nyquist stays zero
        assert np.array_equal(r["acc11"][:-1], r["acc22"][:-1]) and r["acc22"][-1] == 0
        assert np.all(r["acc11"] >= 0)


def test_swapped_stations_conjugate():
    n, fs, nf = 256, 64e6, 64
    r1, r2, _ = make_pair_raw(nf, n, seed=2, delay=3.0)
    cfg = _zero_cfg(fs, n)
    a = R.sector_run(_desc(n, fs), r1, r2, 0, nf, cfg, mode=R.MODE_F64)
    b = R.sector_run(_desc(n, fs), r2, r1, 0, nf, cfg, mode=R.MODE_F64)
    ov = slice(0, n // 2)
    assert np.allclose(a["acc12"][ov], np.conj(b["acc12"][ov]), rtol=1e-10, atol=1e-6)
    assert np.allclose(a["acc11"][ov], b["acc22"][ov]) and np.allclose(a["acc22"][ov], b["acc11"][ov])


def _numpy_sector(raw1, raw2, n, nf, fs, cfg_kwargs, lv=LEVELS_2BIT):
    """Independent FP64 re-run: numpy decode, np.fft.rfft, closed-form phases."""
    codes1 = ((np.frombuffer(raw1.tobytes(), np.uint8)[:, None] >> np.arange(0, 8, 2)) & 3).reshape(nf, n)
    codes2 = ((np.frombuffer(raw2.tobytes(), np.uint8)[:, None] >> np.arange(0, 8, 2)) & 3).reshape(nf, n)
    x1, x2 = lv[codes1], lv[codes2]
    cfg = R.make_delay_cfg(np.zeros(4), **cfg_kwargs)
    half = n // 2 + 1
    acc11 = np.zeros(half)
    acc22 = np.zeros(half)
    acc12 = np.zeros(half, dtype=complex)
    df = fs / n
    k = np.arange(n // 2)
    for i in range(nf):
        d = R.frame_delay(i, cfg)
        s = d.int1
        y1 = np.zeros(n)
        if abs(s) < n:
            if s >= 0:
                y1[s:] = x1[i, :n - s]
            else:
                y1[:n + s] = x1[i, -s:]
        X1 = np.fft.rfft(y1)
        X2 = np.fft.rfft(x2[i])
        acc11 += np.abs(X1) ** 2
        acc22[: n // 2] += np.abs(X2[: n // 2]) ** 2
        ph = np.exp(-2j * np.pi * df * (d.frac1 - d.frac2) * k)
        acc12[: n // 2] += X1[: n // 2] * np.conj(X2[: n // 2]) * ph
    return acc11, acc22, acc12


def test_sector_vs_independent_numpy():
    n, fs, nf = 512, 64e6, 40
    r1, r2, _ = make_pair_raw(nf, n, seed=4, delay=3.3)
    kw = dict(frame_dt=n / fs, fs=fs, time_offset_s=0.0, coarse_delay_s=0.0,
              delay_user_samples=3.3, extra_delay_rate_sps=1e-6, clock1_delay_s=0.0,
              clock1_rate_sps=0.0, clock2_delay_s=0.0, clock2_rate_sps=0.0, d_seek=0.0,
              net_d1_base=0.0, total_rate1_base=0.0, total_accel1_base=0.0, net_d2_base=0.0,
              total_rate2_base=0.0, total_accel2_base=0.0, lo1_hz=0.0, lo2_hz=0.0)
    cfg = R.make_delay_cfg(np.zeros(4), **kw)
    a11, a22, a12 = _numpy_sector(r1, r2, n, nf, fs, kw)
    for mode, tol in ((R.MODE_F64, 1e-11), (R.MODE_F32, 2e-5)):
        r = R.sector_run(_desc(n, fs), r1, r2, 0, nf, cfg, mode=mode)
        assert np.max(np.abs(r["acc11"] - a11) / a11) < tol
        assert np.max(np.abs(r["acc22"][:-1] - a22[:-1]) / a22[:-1]) < tol
        rel = np.abs(r["acc12"][:-1] - a12[:-1]) / np.abs(a12[:-1])
        assert np.median(rel) < tol and np.max(rel) < 50 * tol


@pytest.mark.parametrize("delay", [3.3, 7.75, -2.4])
def test_delay_correction_flattens_phase(delay):
    # after the integer shift and the fractional-delay phase, the XCF phase of
    # a strongly correlated pair is flat: fitted slope << the 2*pi*frac/n slope
    # an uncorrected fractional delay would leave.
    n, fs, nf = 512, 64e6, 100
    r1, r2, _ = make_pair_raw(nf, n, seed=4, delay=delay, rho=0.6)
    cfg = _zero_cfg(fs, n, delay_user_samples=delay)
    r = R.sector_run(_desc(n, fs), r1, r2, 0, nf, cfg, mode=R.MODE_F64)
    k = np.arange(20, 200)
    slope, offset = np.polyfit(k, np.unwrap(np.angle(r["acc12"][20:200])), 1)
    frac = delay - round(delay)
    assert abs(slope) < 0.1 * max(abs(2 * np.pi * frac / n), 2 * np.pi * 0.25 / n)
    assert abs(offset) < 0.1


def test_f32_threads_independent():
    n, fs, nf = 256, 64e6, 700
    r1, r2, _ = make_pair_raw(nf, n, seed=6)
    cfg = _zero_cfg(fs, n)
    a = R.sector_run(_desc(n, fs), r1, r2, 0, nf, cfg, threads=1)
    b = R.sector_run(_desc(n, fs), r1, r2, 0, nf, cfg, threads=4)
    for key in ("acc11", "acc22", "acc12"):
        assert np.array_equal(a[key], b[key])


def test_grid_ant2_and_overlap_only():
    n, fs, nf = 256, 64e6, 32
    r1, r2, _ = make_pair_raw(nf, n, seed=7)
    cfg = _zero_cfg(fs, n)
    d1 = _desc(n, fs, a1s=10, a1e=128, a2s=0)           # station 2 shifted by -10 bins
    g1 = R.sector_run(d1, r1, r2, 0, nf, cfg, mode=R.MODE_F64)
    g2 = R.sector_run(dict(d1, out_grid=R.GRID_ANT2), r1, r2, 0, nf, cfg, mode=R.MODE_F64)
    # same products, indexed on the other station's bins
    assert np.allclose(g1["acc12"][10:128], g2["acc12"][0:118])
    assert np.allclose(g1["acc22"][10:128], g2["acc22"][0:118])
    assert np.all(g1["acc22"][:10] == 0) and np.all(g2["acc11"][118:] == 0)
    oo = R.sector_run(dict(d1, acf_overlap_only=1, need_xcf=0), r1, r2, 0, nf, cfg,
                      mode=R.MODE_F64)
    assert np.all(oo["acc11"][:10] == 0) and np.allclose(oo["acc11"][10:128], g1["acc11"][10:128])
    assert np.all(oo["acc12"] == 0)
