"""Workload descriptions shared by the parity tests: one dict drives both
the oracle (fxref) and the HIP library (fxcorr) on the same seeded bytes."""
from __future__ import annotations

import numpy as np

import fxcorr as F
import fxref as R
from _synth import IDENT, LEVELS_2BIT, make_pair_raw, uniform_levels

ZERO_DELAY = dict(time_offset_s=0.0, coarse_delay_s=0.0, delay_user_samples=0.0,
                  extra_delay_rate_sps=0.0, clock1_delay_s=0.0, clock1_rate_sps=0.0,
                  clock2_delay_s=0.0, clock2_rate_sps=0.0, d_seek=0.0, net_d1_base=0.0,
                  total_rate1_base=0.0, total_accel1_base=0.0, net_d2_base=0.0,
                  total_rate2_base=0.0, total_accel2_base=0.0, lo1_hz=0.0, lo2_hz=0.0)


def make_case(n=1024, fs=64e6, nf=256, bits=(2, 2), shuffles=(None, None), lsb=(False, False),
              band=None, grid=0, products=F.PRODUCT_ACF | F.PRODUCT_XCF, seed=0, rho=0.05,
              delay_samples=5.3, table=None, first_frame=0, **delay_kw):
    sh = tuple(IDENT if s is None else np.asarray(s, dtype=np.int32) for s in shuffles)
    raw1, raw2, levels = make_pair_raw(nf, n, bits=bits, sh=sh, lsb=lsb, rho=rho,
                                       delay=delay_samples, seed=seed)
    if band is None:
        band = (0, 0, n // 2, 0)
    dk = dict(ZERO_DELAY)
    dk.update(frame_dt=n / fs, fs=fs)
    dk.update(delay_kw)
    if table is None and "no_table" not in delay_kw:
        table = np.zeros(max(2, int(np.ceil(nf * n / fs)) + 1))
    dk.pop("no_table", None)
    return dict(n=n, fs=fs, nf=nf, bits=bits, levels=levels, shuffles=sh, lsb=lsb, band=band,
                grid=grid, products=products, raw1=raw1, raw2=raw2, delay=dk, table=table,
                first_frame=first_frame)


def oracle_run(c, mode=R.MODE_F32, threads=8):
    n = c["n"]
    d = dict(fft_len=n, bits1=c["bits"][0], bits2=c["bits"][1], levels1=c["levels"][0],
             levels2=c["levels"][1], shuffle1=c["shuffles"][0], shuffle2=c["shuffles"][1],
             lsb1=int(c["lsb"][0]), lsb2=int(c["lsb"][1]), a1s=c["band"][1], a1e=c["band"][2],
             a2s=c["band"][3], out_grid=c["grid"], df_hz=c["fs"] / n,
             need_acf=int(bool(c["products"] & F.PRODUCT_ACF)),
             need_xcf=int(bool(c["products"] & F.PRODUCT_XCF)),
             acf_overlap_only=int(bool(c["products"] & F.PRODUCT_ACF_OVERLAP_ONLY)),
             cpu_threads=threads)
    cfg = R.make_delay_cfg(c["table"], **c["delay"])
    return R.sector_run(d, c["raw1"], c["raw2"], c["first_frame"], c["nf"], cfg, mode=mode,
                        threads=threads)


def gpu_plan(c, precision=F.PRECISION_PLAIN_FP32):
    return F.Plan(c["n"], bits=c["bits"], levels=c["levels"], shuffles=c["shuffles"],
                  lsb=c["lsb"], band=c["band"], df_hz=c["fs"] / c["n"], products=c["products"],
                  out_grid=c["grid"], precision=precision)


def gpu_run(c, precision=F.PRECISION_PLAIN_FP32, plan=None):
    p = plan or gpu_plan(c, precision)
    cfg = F.make_delay_cfg(c["table"], **c["delay"])
    return p.sector_run(c["raw1"], c["raw2"], c["first_frame"], c["nf"], cfg)


def rel_err(got, ref):
    """Per-bin relative error |got-ref|/|ref| over bins where ref != 0."""
    got = np.asarray(got)
    ref = np.asarray(ref)
    m = np.abs(ref) > 0
    if not np.any(m):
        return 0.0
    return float(np.max(np.abs(got[m] - ref[m]) / np.abs(ref[m])))


def xcf_err(got, ref):
    """Relative error of a cross spectrum.  Per-bin |got-ref|/|ref| on bins at
    least 0.1x the spectrum's RMS magnitude, and |got-ref|/RMS on every bin:
    in a bin where the frame sum nearly cancels, per-frame FP32 rounding noise
    (uncorrelated across frames) divided by the near-zero sum says nothing
    about the arithmetic, so those bins are held to the RMS scale instead."""
    got = np.asarray(got)
    ref = np.asarray(ref)
    m = np.abs(ref) > 0
    if not np.any(m):
        return 0.0
    rms = float(np.sqrt(np.mean(np.abs(ref[m]) ** 2)))
    strong = np.abs(ref) >= 0.1 * rms
    e_strong = float(np.max(np.abs(got[strong] - ref[strong]) / np.abs(ref[strong])))
    e_all = float(np.max(np.abs(got - ref)) / rms)
    return max(e_strong, e_all)


def spectra_err(got, ref):
    """{acc11, acc22: per-bin relative; acc12: xcf_err} for result dicts."""
    return {"acc11": rel_err(got["acc11"], ref["acc11"]),
            "acc22": rel_err(got["acc22"], ref["acc22"]),
            "acc12": xcf_err(got["acc12"], ref["acc12"])}


def amp_phase_err(got, ref):
    """Per-channel amplitude and phase error of a spectrum against a reference
    spectrum (the products written to .cor, main.rs:2660-2683).  Over channels
    with ref != 0: amplitude error ||got|-|ref||/|ref| and phase error
    |arg(got * conj(ref))| in radians, as max and RMS; "_strong" repeats the
    max over channels at least 0.1x the reference's RMS magnitude (a channel
    whose frame sum nearly cancels magnifies rounding noise that is
    uncorrelated across frames)."""
    got = np.asarray(got).astype(np.complex128)
    ref = np.asarray(ref).astype(np.complex128)
    m = np.abs(ref) > 0
    if not np.any(m):
        return dict(amp_max=0.0, amp_rms=0.0, phase_max=0.0, phase_rms=0.0, amp_max_strong=0.0,
                    phase_max_strong=0.0, channels=0)
    g, r = got[m], ref[m]
    amp = np.abs(np.abs(g) - np.abs(r)) / np.abs(r)
    ph = np.abs(np.angle(g * np.conj(r)))
    rms = np.sqrt(np.mean(np.abs(r) ** 2))
    s = np.abs(r) >= 0.1 * rms
    return dict(amp_max=float(amp.max()), amp_rms=float(np.sqrt(np.mean(amp ** 2))),
                phase_max=float(ph.max()), phase_rms=float(np.sqrt(np.mean(ph ** 2))),
                amp_max_strong=float(amp[s].max()), phase_max_strong=float(ph[s].max()),
                channels=int(m.sum()))


def spectra_amp_phase(got, ref):
    """amp_phase_err per product of two result dicts (acc11, acc22, acc12)."""
    return {k: amp_phase_err(got[k], ref[k]) for k in ("acc11", "acc22", "acc12")}


def fmt_amp_phase(d):
    """One line per product: amplitude max/rms (relative), phase max/rms (rad)."""
    return "; ".join(f"{k}: amp {v['amp_max']:.2e}/{v['amp_rms']:.2e} "
                     f"phase {v['phase_max']:.2e}/{v['phase_rms']:.2e} rad"
                     for k, v in d.items())


def zero_pattern_equal(got, ref):
    return np.array_equal(np.asarray(got) == 0, np.asarray(ref) == 0)


def out_decode_plan(bit_out, shuffle1, levels1):
    """Decoder of phased .raw words: bit_out-bit codes under station 1's
    shuffle, valued by station 1's levels (main.rs:2295).  When bit_out
    exceeds station 1's width, codes past its levels never occur; they are
    given distinct sentinel values so the table has 2^bit_out entries."""
    lv = np.asarray(levels1, dtype=np.float64)
    full = np.concatenate([lv, 1e6 + np.arange((1 << bit_out) - len(lv), dtype=np.float64)])
    return R.DecodePlan(bit_out, shuffle1, full)
