"""Precision stress at its own workload, through the C ABI: 512 Msps, station A
4-bit (16 uniform levels) / station B 1-bit, FFT 65536, 16 MHz centre offset
(band shift 2048 bins, overlap 30,720; main.rs:360-372), a correlated
continuum of coherence 1e-3 plus a Gaussian spectral line three channels wide
(FWHM) of peak coherence 0.05 at a seeded channel, 30 s = 234,375 frames in 30
sectors of 7,812/7,813 frames (main.rs:2058-2072).

- One full 7,812-frame sector, generated on the device and downloaded: the
  plain form against the reference-faithful FP32 oracle (1e-5), the chosen
  form against the FP64 oracle (1e-4), per-channel amplitude and phase
  errors printed for both.
- All 30 sectors on the device (too large for the oracle), checked by what
  the signal fixes: the line is the strongest channel of the normalised
  cross spectrum, with the expected coherence; the continuum's mean
  coherence is the expected one; ACFs are positive; the chosen and plain
  forms agree within the chosen form's bar."""
import os

import numpy as np
import pytest

import fxcorr as F
import fxref as R
from _cases import fmt_amp_phase, spectra_amp_phase, spectra_err, zero_pattern_equal
from _synth import uniform_levels

pytestmark = pytest.mark.gpu

N, FS = 65536, 512e6
BITS = (4, 1)
LEVELS = (uniform_levels(4), uniform_levels(1))
TOTAL_FRAMES = int(30 * FS // N)                 # 234,375
RHO_C, RHO_L, FWHM = 1e-3, 0.05, 3.0
SEED = 2303
LINE = int(np.random.default_rng(SEED).integers(2000, 28000))   # station-1 channel
PLAIN_TOL, CHOSEN_TOL = 1e-5, 1e-4


def _band():
    band = F.band_alignment(N, FS, 8400.0, 8384.0, 256.0, 256.0)
    assert band == (2048, 0, 30720, 2048)
    return band


def _generate(first, nf, stream):
    d1 = F.DeviceBuffer(nf * N * BITS[0] // 8)
    d2 = F.DeviceBuffer(nf * N * BITS[1] // 8)
    F.synth_spectral_pair_device(d1.ptr, d2.ptr, N, first, nf, bits=BITS, shift_bins=2048,
                                 rho_continuum=RHO_C, rho_line=RHO_L, line_bin=LINE,
                                 line_fwhm_bins=FWHM, clip_sigma=4.0, seed=SEED,
                                 stream=stream.ptr)
    return d1, d2


def _plan(precision):
    band = _band()
    return F.Plan(N, bits=BITS, levels=list(LEVELS), band=band, df_hz=FS / N,
                  products=F.PRODUCT_ACF | F.PRODUCT_XCF, precision=precision)


def _delay(**kw):
    d = dict(frame_dt=N / FS, fs=FS, time_offset_s=0.0, coarse_delay_s=0.0,
             delay_user_samples=0.0, extra_delay_rate_sps=0.0, clock1_delay_s=0.0,
             clock1_rate_sps=0.0, clock2_delay_s=0.0, clock2_rate_sps=0.0, d_seek=0.0,
             net_d1_base=0.0, total_rate1_base=0.0, total_accel1_base=0.0, net_d2_base=0.0,
             total_rate2_base=0.0, total_accel2_base=0.0, lo1_hz=0.0, lo2_hz=0.0)
    d.update(kw)
    return d


def _run(plan, d1, d2, first, nf, dcfg, stream):
    acc = F.DeviceBuffer(F.N_PRODUCTS * plan.half * 8)
    plan.sector_run_device(d1, d2, first, nf, F.make_delay_cfg(None, **dcfg), acc.ptr,
                           stream=stream.ptr)
    stream.sync()
    F.plan_check_device_error(plan, stream)
    a = acc.download(np.zeros(F.N_PRODUCTS * plan.half)).reshape(F.N_PRODUCTS, plan.half)
    acc.free()
    return dict(acc11=a[1], acc22=a[2], acc12=a[3] + 1j * a[4])


def _quant_efficiency():
    """Coherence of the quantised pair over that of the Gaussians at small
    coherence: E[x q1(x)] E[y q2(y)] / sqrt(E[q1^2] E[q2^2]) (Bussgang)."""
    x = np.linspace(-12, 12, 2_000_001)
    w = np.exp(-x * x / 2) / np.sqrt(2 * np.pi)
    dx = x[1] - x[0]
    q4 = np.clip(np.floor(x / 0.5 + 8), 0, 15) - 7.5
    q1 = np.where(x < 0, -0.5, 0.5)
    e = [(np.sum(x * q * w) * dx, np.sum(q * q * w) * dx) for q in (q4, q1)]
    return e[0][0] * e[1][0] / np.sqrt(e[0][1] * e[1][1])


def test_c3_one_sector_vs_oracle():
    sectors = F.sector_partition(TOTAL_FRAMES, N / FS)
    assert len(sectors) == 30 and sorted(set(sectors)) == [7812, 7813]
    si = 15                                                    # first 7,812-frame sector
    first, nf = int(sum(sectors[:si])), int(sectors[si])
    assert nf == 7812
    print(f"[c3] sector {si}: frames {first}..{first + nf - 1}, line at channel {LINE}", flush=True)
    s = F.Stream()
    d1, d2 = _generate(first, nf, s)
    h1 = d1.download(np.zeros(nf * N * BITS[0] // 8, dtype=np.uint8))
    h2 = d2.download(np.zeros(nf * N * BITS[1] // 8, dtype=np.uint8))
    # a 1.7-sample delay (integer shift 1 plus a fraction), 0.3 samples of
    # station-1 clock and a rate small enough (0.05 rad of fringe over the
    # sector) that the generated, undelayed line stays coherent
    dcfg = _delay(delay_user_samples=1.7, extra_delay_rate_sps=1e-12, lo1_hz=8.4e9,
                  lo2_hz=8.384e9, clock1_delay_s=0.3 / FS)
    band = _band()
    threads = max(1, min(os.cpu_count() or 1, 16))
    desc = dict(fft_len=N, bits1=BITS[0], bits2=BITS[1], levels1=LEVELS[0], levels2=LEVELS[1],
                a1s=band[1], a1e=band[2], a2s=band[3], out_grid=0, df_hz=FS / N, need_acf=1,
                need_xcf=1, cpu_threads=threads)
    rcfg = R.make_delay_cfg(None, **dcfg)
    res = {}
    for name, prec in (("plain", F.PRECISION_PLAIN_FP32), ("chosen", F.PRECISION_CHOSEN)):
        p = _plan(prec)
        res[name] = _run(p, d1.ptr, d2.ptr, first, nf, dcfg, s)
        p.close()
        print(f"[c3] {name} form done", flush=True)
    d1.free()
    d2.free()
    o32 = R.sector_run(desc, h1, h2, first, nf, rcfg, mode=R.MODE_F32, threads=threads)
    print("[c3] oracle FP32 done", flush=True)
    o64 = R.sector_run(desc, h1, h2, first, nf, rcfg, mode=R.MODE_F64, threads=threads)
    print("[c3] oracle FP64 done", flush=True)
    for name, ref, tol in (("plain", o32, PLAIN_TOL), ("chosen", o64, CHOSEN_TOL)):
        got = res[name]
        e = spectra_err(got, ref)
        ap = spectra_amp_phase(got, ref)
        print(f"[c3] {name} vs oracle: {e}; per channel: {fmt_amp_phase(ap)}", flush=True)
        for k in ("acc11", "acc22", "acc12"):
            assert zero_pattern_equal(got[k], ref[k]), (name, k)
            assert e[k] < tol, (name, k, e[k])
            assert ap[k]["amp_max_strong"] < tol and ap[k]["phase_max_strong"] < tol, (name, k)
    # the oracle itself sees the generated line where it was put, at the
    # expected coherence (one sector: noise 1/sqrt(2 nf) per component)
    # (on station 1's grid, ACF2 and the XCF sit at station-1 channels a1s..a1e)
    x = o64["acc12"][band[1]:band[2]]
    coh = np.abs(x) / np.sqrt(o64["acc11"][band[1]:band[2]] * o64["acc22"][band[1]:band[2]])
    sigma = 1.0 / np.sqrt(2.0 * nf)
    print(f"[c3] oracle: line coherence {coh[LINE - band[1]]:.4f} (expected "
          f"{RHO_L * _quant_efficiency():.4f}, sigma {sigma:.4f})", flush=True)
    assert abs(coh[LINE - band[1]] - RHO_L * _quant_efficiency()) < 4.5 * sigma


def test_c3_thirty_sectors_properties():
    sectors = F.sector_partition(TOTAL_FRAMES, N / FS)
    s = F.Stream()
    print(f"[c3] generating {TOTAL_FRAMES} frames (line at channel {LINE})", flush=True)
    d1, d2 = _generate(0, TOTAL_FRAMES, s)
    bpf1, bpf2 = N * BITS[0] // 8, N * BITS[1] // 8
    dcfg = _delay()
    band = _band()
    tot = {}
    for name, prec in (("plain", F.PRECISION_PLAIN_FP32), ("chosen", F.PRECISION_CHOSEN)):
        p = _plan(prec)
        acc = F.DeviceBuffer(len(sectors) * F.N_PRODUCTS * p.half * 8)
        cfg = F.make_delay_cfg(None, **dcfg)
        first = 0
        for i, nf in enumerate(sectors):
            p.sector_run_device(d1.ptr + first * bpf1, d2.ptr + first * bpf2, first, nf, cfg,
                                acc.ptr + i * F.N_PRODUCTS * p.half * 8, stream=s.ptr)
            first += nf
        s.sync()
        F.plan_check_device_error(p, s)
        a = acc.download(np.zeros(len(sectors) * F.N_PRODUCTS * p.half))
        a = a.reshape(len(sectors), F.N_PRODUCTS, p.half)
        acc.free()
        p.close()
        assert np.all(np.isfinite(a))
        # every sector's ACFs: station 1 on all its channels, station 2 on the
        # overlap mapped to station 1's grid (zero outside it, main.rs:2492-2529)
        assert np.all(a[:, 1] > 0), name
        assert np.all(a[:, 2, band[1]:band[2]] > 0) and np.all(a[:, 2, band[2]:] == 0), name
        t = a.sum(axis=0)
        tot[name] = dict(acc11=t[1], acc22=t[2], acc12=t[3] + 1j * t[4])
        print(f"[c3] {name}: 30 sectors done", flush=True)
    d1.free()
    d2.free()
    eta = _quant_efficiency()
    for name, t in tot.items():
        coh = t["acc12"][band[1]:band[2]] / np.sqrt(t["acc11"][band[1]:band[2]] *
                                                     t["acc22"][band[1]:band[2]])
        peak = int(np.argmax(np.abs(coh))) + band[1]
        sigma = 1.0 / np.sqrt(2.0 * TOTAL_FRAMES)                 # per real part, per channel
        line = coh[LINE - band[1]]
        far = np.abs(np.arange(band[1], band[2]) - LINE) > 20
        cont = float(np.mean(coh[far].real))
        print(f"[c3] {name}: peak channel {peak} (line {LINE}), line coherence {line.real:.5f}"
              f"{line.imag:+.5f}j (expected {RHO_L * eta:.5f}), continuum {cont:.3e} "
              f"(expected {RHO_C * eta:.3e}), efficiency {eta:.4f}", flush=True)
        assert peak == LINE, name
        assert abs(line.real - RHO_L * eta) < 6 * sigma and abs(line.imag) < 6 * sigma, name
        assert abs(cont - RHO_C * eta) < 6 * sigma / np.sqrt(far.sum()) + 0.02 * RHO_C, name
        assert abs(float(np.mean(coh[far].imag))) < 6 * sigma / np.sqrt(far.sum()), name
    e = spectra_err(tot["chosen"], tot["plain"])
    ap = spectra_amp_phase(tot["chosen"], tot["plain"])
    print(f"[c3] chosen vs plain over 30 sectors: {e}; per channel: {fmt_amp_phase(ap)}",
          flush=True)
    assert all(v < CHOSEN_TOL for v in e.values()), e
