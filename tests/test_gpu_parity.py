"""GPU parity of the HIP hot path against the oracle, called through the C ABI.

Bars: decode, integer shifts and delay integers bit-exact;
plain FP32 spectra within 1e-5 relative of the reference-faithful oracle;
chosen form within 1e-4 relative of the FP64 oracle."""
import numpy as np
import pytest

import fxcorr as F
import fxref as R
from _cases import (gpu_plan, gpu_run, make_case, oracle_run, out_decode_plan, rel_err,
                    spectra_err, zero_pattern_equal, spectra_amp_phase,
                    fmt_amp_phase)
from _synth import IDENT, LEVELS_2BIT, uniform_levels

pytestmark = pytest.mark.gpu

PLAIN_TOL = 1e-5
CHOSEN_TOL = 1e-4


def _plan(n, bits=(2, 2), levels=None, shuffles=(None, None), lsb=(False, False), **kw):
    if levels is None:
        levels = [LEVELS_2BIT if b == 2 else uniform_levels(b) for b in bits]
    return F.Plan(n, bits=bits, levels=levels, shuffles=shuffles, lsb=lsb, df_hz=1.0, **kw)


# ----------------------------------------------------------------- decode
@pytest.mark.parametrize("bits", [1, 2, 3, 4, 5, 8, 12, 16])
@pytest.mark.parametrize("kind", ["ident", "pair", "generic"])
def test_decode_bit_exact(bits, kind):
    rng = np.random.default_rng(100 * bits + len(kind))
    n, nfr = 1024, 24
    sh = {"ident": IDENT, "pair": IDENT ^ 1, "generic": rng.permutation(32).astype(np.int32)}[kind]
    lv = rng.standard_normal(1 << bits)
    bpf = n * bits // 8
    raw = rng.integers(0, 256, size=nfr * bpf, dtype=np.uint8)
    shifts = rng.integers(-40, 40, size=nfr)
    shifts[:4] = [0, n, -n, 5000]
    for lsb in (False, True):
        p = _plan(n, bits=(bits, 2), levels=[lv, LEVELS_2BIT], shuffles=(sh, None),
                  lsb=(lsb, False))
        got, nfail = p.decode_frames(0, raw, nfr, shifts)
        assert nfail == 0
        op = R.DecodePlan(bits, sh, lv)
        for f in range(nfr):
            ok, v = op.decode(raw[f * bpf:(f + 1) * bpf], n, lsb)
            assert ok
            exp = R.shift_zerofill(v, int(shifts[f]))
            assert np.array_equal(got[f].view(np.uint32), exp.view(np.uint32)), (f, lsb)
        p.close()


def test_decode_station2_and_no_shift():
    rng = np.random.default_rng(9)
    n, nfr = 256, 10
    sh = rng.permutation(32).astype(np.int32)
    raw = rng.integers(0, 256, size=nfr * n // 4, dtype=np.uint8)
    p = _plan(n, shuffles=(None, sh), lsb=(False, True))
    got, _ = p.decode_frames(1, raw, nfr)
    op = R.DecodePlan(2, sh, LEVELS_2BIT)
    for f in range(nfr):
        ok, v = op.decode(raw[f * 64:(f + 1) * 64], n, True)
        assert ok and np.array_equal(got[f], v)


# ------------------------------------------------------------------ delay
def _delay_cfg_pair(table, **kw):
    base = dict(frame_dt=1024 / 1024e6, fs=1024e6, time_offset_s=3.0, coarse_delay_s=1.1e-9,
                delay_user_samples=0.2, extra_delay_rate_sps=7.46e-12, clock1_delay_s=0.0,
                clock1_rate_sps=0.0, clock2_delay_s=1.2e-6, clock2_rate_sps=1e-12,
                d_seek=1536 / 1024e6, net_d1_base=3.1e-7, total_rate1_base=2.2e-11,
                total_accel1_base=1e-15, net_d2_base=0.0, total_rate2_base=0.0,
                total_accel2_base=0.0, lo1_hz=6.7e9, lo2_hz=6.7e9)
    base.update(kw)
    return R.make_delay_cfg(table, **base), F.make_delay_cfg(table, **base)


@pytest.mark.parametrize("with_table", [True, False])
def test_delay_eval_bit_exact(with_table):
    table = (1.5e-7 + 3e-9 * np.sin(np.arange(10) / 3.0)) if with_table else None
    rc, fc = _delay_cfg_pair(table)
    idx = np.concatenate([np.arange(0, 2000), np.arange(999_000, 1_001_000),
                          np.arange(7_999_000, 8_000_000)])
    got = F.delay_eval(fc, 0, int(idx.max()) + 1)
    ulp_bad = 0
    for i in idx[::7]:
        g, o = got[int(i)], R.frame_delay(int(i), rc)
        assert g.int1 == o.int1 and g.int2 == o.int2
        assert g.frac1 == o.frac1 and g.frac2 == o.frac2
        for a, b in ((g.fr_lo1, o.fr_lo1), (g.fr_lo2, o.fr_lo2)):
            ga = np.array(a[:], dtype=np.float32).view(np.int32)
            ob = np.array(b[:], dtype=np.float32).view(np.int32)
            d = np.abs(ga.astype(np.int64) - ob.astype(np.int64))
            # f32 cos/sin: the reference's glibc cosf/sinf are not correctly
            # rounded in ~1% of arguments; the device returns the correctly
            # rounded value, so at most 1 ulp apart, rarely.
            assert np.all(d <= 1)
            ulp_bad += int(np.sum(d))
    assert ulp_bad <= 4 * len(idx[::7]) // 50


# -------------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384])
def test_rfft_vs_numpy_and_oracle(n):
    p = _plan(n)
    x = np.random.default_rng(n).standard_normal((5, n)).astype(np.float32)
    got = p.rfft_frames(x)
    ref = np.fft.rfft(x.astype(np.float64), axis=1)
    scale = np.sqrt(n)
    err = np.sqrt(np.mean(np.abs(got - ref) ** 2, axis=1)) / scale
    assert np.all(err < 3e-7 * np.sqrt(np.log2(n)))
    assert np.all(got[:, 0].imag == 0) and np.all(got[:, -1].imag == 0)
    orc = np.stack([R.rfft_f32(r) for r in x])
    assert np.max(np.abs(got - orc)) / scale < 2e-6
    y = p.irfft_frames(got)
    assert np.max(np.abs(y - x)) < 1e-5 * np.sqrt(np.log2(n))
    p.close()


def test_irfft_matches_oracle():
    n = 1024
    p = _plan(n)
    rng = np.random.default_rng(4)
    spec = (rng.standard_normal((3, n // 2 + 1)) + 1j * rng.standard_normal((3, n // 2 + 1)))
    spec = spec.astype(np.complex64)
    got = p.irfft_frames(spec)
    for i in range(3):
        assert np.allclose(got[i], R.irfft_f32(spec[i], n), atol=2e-6)


@pytest.mark.parametrize("n", [32768, 65536])
def test_irfft_large_matches_oracle(n):
    """inverse_c2r_process through the HBM four-step (the .raw path of the
    phased sum at N >= 32768)."""
    p = _plan(n)
    assert p.info.fft_form == 2
    rng = np.random.default_rng(n + 4)
    spec = (rng.standard_normal((3, n // 2 + 1)) + 1j * rng.standard_normal((3, n // 2 + 1)))
    spec = spec.astype(np.complex64)
    got = p.irfft_frames(spec)
    for i in range(3):
        ref = R.irfft_f32(spec[i], n)
        rms = np.sqrt(np.mean(ref.astype(np.float64) ** 2))
        assert np.max(np.abs(got[i] - ref)) < 2e-5 * rms, i


# --------------------------------------------------------------- quantise
@pytest.mark.parametrize("bits", [1, 2, 4, 8])
def test_quantise_bit_exact(bits):
    rng = np.random.default_rng(bits + 40)
    n = 1024
    lv = uniform_levels(bits) * 0.9
    sh = rng.permutation(32).astype(np.int32)
    p = F.Plan(n, bits=(bits, bits), levels=[lv, lv], shuffles=(sh, None), df_hz=1.0,
               products=F.PRODUCT_ACF, bit_out=bits)
    x = (rng.standard_normal((6, n)) * lv.max()).astype(np.float32)
    x[0, :5] = [np.nan, np.inf, -np.inf, 0.0, lv[0]]
    got = p.quantise_frames(x)
    for f in range(6):
        assert np.array_equal(got[f], R.quantise_frame(x[f], bits, lv, sh))
    p.close()


@pytest.mark.parametrize("kind", ["uniform", "near_uniform", "gaussian", "duplicates", "unsorted"])
def test_quantise_nearest_level_edges(kind):
    """Evenly spaced levels (exactly, or within a quarter step) take the
    spacing guess and its four-level window, other ascending levels the
    bisection, others the linear scan; all give the reference's first
    minimum of |x - level| (f32) bit for bit at midpoints (ties go to the
    lower index), at the levels, one ulp either side of the midpoints,
    beyond the outer levels and for NaN / +-inf."""
    rng = np.random.default_rng(len(kind))
    bits, n = 8, 1024
    jitter = rng.uniform(-0.12, 0.12, 256)
    jitter[[0, -1]] = 0.0                               # within a quarter step of the end-to-end line
    lv = {"uniform": uniform_levels(8) * 0.5,
          "near_uniform": (np.arange(256) - 127.5 + jitter) * 0.37,
          "gaussian": np.sort(rng.standard_normal(256)) * 3.0,
          "duplicates": np.repeat(np.sort(rng.standard_normal(64)), 4),
          "unsorted": rng.standard_normal(256) * 2.0}[kind]
    lv32 = lv.astype(np.float32)
    srt = np.sort(lv32).astype(np.float64)
    mids = ((srt[:-1] + srt[1:]) / 2).astype(np.float32)
    pool = np.concatenate([mids, lv32, np.nextafter(mids, np.float32(np.inf)),
                           np.nextafter(mids, np.float32(-np.inf)),
                           np.float32([np.nan, np.inf, -np.inf, 1e30, -1e30, 0.0, -0.0]),
                           (rng.standard_normal(4000) * 4).astype(np.float32)])
    x = rng.choice(pool, size=(6, n)).astype(np.float32)
    sh = np.arange(32, dtype=np.int32)                  # identity output bit map
    p = F.Plan(n, bits=(bits, bits), levels=[lv, lv], shuffles=(sh, None), df_hz=1.0,
               products=F.PRODUCT_ACF, bit_out=bits)
    got = p.quantise_frames(x)
    for f in range(6):
        assert np.array_equal(got[f], R.quantise_frame(x[f], bits, lv, sh)), f
    p.close()


@pytest.mark.parametrize("bits", [2, 3, 4])
def test_quantise_near_uniform_narrow_codes(bits):
    """Narrow output codes with levels within a quarter step of even spacing
    (the spacing guess with a window that covers most or all of the levels):
    bit for bit the reference's first minimum at midpoints, levels, one ulp
    either side of the midpoints, beyond the outer levels and NaN / +-inf."""
    rng = np.random.default_rng(bits + 90)
    nl, n = 1 << bits, 1024
    jitter = rng.uniform(-0.12, 0.12, nl)
    jitter[[0, -1]] = 0.0
    lv = ((np.arange(nl) - (nl - 1) / 2 + jitter) * 0.8).astype(np.float32)
    srt = lv.astype(np.float64)
    mids = ((srt[:-1] + srt[1:]) / 2).astype(np.float32)
    pool = np.concatenate([mids, lv, np.nextafter(mids, np.float32(np.inf)),
                           np.nextafter(mids, np.float32(-np.inf)),
                           np.float32([np.nan, np.inf, -np.inf, 1e30, -1e30, 0.0, -0.0]),
                           (rng.standard_normal(2000) * nl * 0.5).astype(np.float32)])
    x = rng.choice(pool, size=(4, n)).astype(np.float32)
    sh = rng.permutation(32).astype(np.int32)
    p = F.Plan(n, bits=(bits, bits), levels=[lv, lv], shuffles=(sh, None), df_hz=1.0,
               products=F.PRODUCT_ACF, bit_out=bits)
    got = p.quantise_frames(x)
    for f in range(4):
        assert np.array_equal(got[f], R.quantise_frame(x[f], bits, lv, sh)), f
    p.close()


# ------------------------------------------------------------- sector MAC
def _compare(got, ref, tol, keys=("acc11", "acc22", "acc12")):
    # ACF: per-bin relative error.  XCF: per-bin relative on bins >= 0.1x the
    # spectrum RMS plus |error| <= tol x RMS everywhere (see _cases.xcf_err).
    # Per-channel amplitude and phase errors are printed for every comparison
    # and held to the same bar on channels >= 0.1x the RMS magnitude.
    errs = spectra_err(got, ref)
    ap = spectra_amp_phase(got, ref)
    print(f"[parity] tol {tol:.0e}: {fmt_amp_phase(ap)}", flush=True)
    for k in keys:
        assert zero_pattern_equal(got[k], ref[k]), k
        assert errs[k] < tol, (k, errs[k])
        assert ap[k]["amp_max_strong"] < tol and ap[k]["phase_max_strong"] < tol, (k, ap[k])


def test_64msps_2bit_fft1024_full_plain_and_chosen():
    """64 Msps 2-bit two-station case in full: N=1024, 62,499 frames, USB,
    37.3-sample delay (seek 32 -> int1 = 5, frac1 = 0.3), fringe at 6.7 GHz."""
    fs, n = 64e6, 1024
    c = make_case(n=n, fs=fs, nf=62499, rho=0.02, delay_samples=37.3 - 32, seed=0,
                  delay_user_samples=37.3, d_seek=32 / fs, extra_delay_rate_sps=0.5 / 6.7e9,
                  lo1_hz=6.7e9, lo2_hz=6.7e9)
    g = gpu_run(c, F.PRECISION_PLAIN_FP32)
    o32 = oracle_run(c, R.MODE_F32)
    assert g["n_failed"] == 0 and o32["n_failed"] == 0
    _compare(g, o32, PLAIN_TOL)
    gc = gpu_run(c, F.PRECISION_CHOSEN)
    o64 = oracle_run(c, R.MODE_F64)
    _compare(gc, o64, CHOSEN_TOL)


def _check_plain(c, g, o32=None):
    """Plain form vs the reference-faithful oracle within 1e-5 at any overlap
    length: beyond 512 overlap bins the device replays the reference's own
    sequential f32 XCF phase chain (main.rs:2526), bit for bit; up to 512
    bins it takes FP64 powers of the same f32 step, which stay within that
    chain's rounding walk (<= 1.4e-5 over a frame, 2.8e-6 rms)."""
    o32 = o32 if o32 is not None else oracle_run(c, R.MODE_F32)
    _compare(g, o32, PLAIN_TOL)


@pytest.mark.parametrize("n", [64, 256, 512, 1024, 2048, 4096, 8192, 16384])
def test_sector_sizes(n):
    nf = max(128, 2 ** 19 // n)
    fs = 256e6
    c = make_case(n=n, fs=fs, nf=nf, rho=0.5, delay_samples=2.6, seed=n,
                  delay_user_samples=2.6, extra_delay_rate_sps=3e-9, lo1_hz=1.0e9, lo2_hz=1.0e9)
    o64 = oracle_run(c, R.MODE_F64)
    _check_plain(c, gpu_run(c))
    _compare(gpu_run(c, F.PRECISION_CHOSEN), o64, CHOSEN_TOL)


def test_1024msps_2bit_fft1024_lsb_shuffle_geometry():
    """1024 Msps 2-bit N=1024 shape: B LSB with a seeded shuffle, geometric
    table + clock 1.2 us + rate 1e-12 + residuals; a 20000-frame slice of
    sector 3 (first_frame != 0).  The table is the library's own 1-s
    geometric table for the reference's default stations (geom.rs:7-8)
    toward a J2000 source precessed to the epoch (main.rs:1128-1135)."""
    fs, n = 1024e6, 1024
    rng = np.random.default_rng(2)
    sh = rng.permutation(32).astype(np.int32)
    mjd = F.parse_epoch_mjd("2024/043 15:52:00")
    ra, dec = F.precess_j2000_to_date(F.parse_ra("16h42m58.8s"), F.parse_dec("+39d48m36.0s"), mjd)
    table = F.geometric_delay_table([-3502544.587, 3950966.235, 3566381.192],
                                    [-3502567.576, 3950885.734, 3566449.115], ra, dec, mjd, 0.0,
                                    F.geometric_table_points(8_000_000, n / fs))
    assert table.size == 9 and np.all(np.abs(table) < 107.81 / 299792458.0)
    init = int(round((table[0] + 0.2 / fs + 1.2e-6) * fs))           # main.rs:1279
    d_seek = (max(init, 0) // 16 * 16) / fs                            # 2-bit word-aligned seek
    c = make_case(n=n, fs=fs, nf=20000, shuffles=(None, sh), lsb=(False, True), rho=0.005,
                  delay_samples=3.0, seed=3, table=table, first_frame=3_000_000,
                  time_offset_s=0.0, clock2_delay_s=1.2e-6, clock2_rate_sps=1e-12,
                  delay_user_samples=0.2, extra_delay_rate_sps=0.05 / 6.7e9,
                  d_seek=d_seek, lo1_hz=6.7e9, lo2_hz=6.7e9)
    _compare(gpu_run(c), oracle_run(c), PLAIN_TOL)
    _compare(gpu_run(c, F.PRECISION_CHOSEN), oracle_run(c, R.MODE_F64), CHOSEN_TOL)


@pytest.mark.parametrize("grid", [0, 1])
@pytest.mark.parametrize("overlap_only", [False, True])
def test_band_shift_grids(grid, overlap_only):
    """Mixed bits (4/1) with a band shift, scaled to N=4096."""
    n, fs = 4096, 512e6
    sb = -128
    band = (sb, 128, 2048, 0) if grid == 0 else (-sb, 0, 2048 - 128, 128)
    prods = F.PRODUCT_ACF | F.PRODUCT_XCF
    if overlap_only:
        prods = F.PRODUCT_ACF | F.PRODUCT_ACF_OVERLAP_ONLY
    c = make_case(n=n, fs=fs, nf=400, bits=(4, 1), band=band, grid=grid, products=prods,
                  rho=0.5, delay_samples=1.4, seed=11, delay_user_samples=1.4,
                  lo1_hz=8.4e9, lo2_hz=8.4e9)
    o64 = oracle_run(c, R.MODE_F64)
    _check_plain(c, gpu_run(c))
    _compare(gpu_run(c, F.PRECISION_CHOSEN), o64, CHOSEN_TOL)


@pytest.mark.parametrize("n", [1024, 2048])
def test_stockham_mfma_band_shift(n):
    """A band shift keeps N = 1024/2048 off the fused MFMA kernels: the
    chosen form runs the Stockham kernel with its MFMA first pass."""
    fs = 512e6
    c = make_case(n=n, fs=fs, nf=1500, bits=(2, 4), band=(-64, 64, n // 2, 0), rho=0.5,
                  delay_samples=-2.3, seed=19, delay_user_samples=-2.3,
                  extra_delay_rate_sps=2e-9, lo1_hz=8.4e9, lo2_hz=8.4e9)
    p = gpu_plan(c, F.PRECISION_CHOSEN)
    assert p.info.fft_form == 0 and p.info.fft_mfma == 1
    o64 = oracle_run(c, R.MODE_F64)
    got = gpu_run(c, plan=p)
    _compare(got, o64, CHOSEN_TOL)
    for k in ("acc11", "acc22", "acc12"):
        assert zero_pattern_equal(got[k], o64[k]), k


def test_analytic_delay_branch_and_large_shifts():
    # no geometric table: quadratic delay model; shifts up to beyond +-N
    n, fs = 512, 128e6
    c = make_case(n=n, fs=fs, nf=600, no_table=True, net_d1_base=-6.0e-6,
                  total_rate1_base=1.2e-2, total_accel1_base=1e-3, lo1_hz=2e9, seed=5)
    d = R.make_delay_cfg(None, **c["delay"])
    ints = [R.frame_delay(i, d).int1 for i in range(600)]
    assert min(ints) < -n and max(ints) > 0   # frames fully zeroed and shifted both ways
    _compare(gpu_run(c), oracle_run(c), PLAIN_TOL)


@pytest.mark.parametrize("bits", [(1, 8), (4, 2), (2, 2)])
def test_stockham_mfma_large_and_odd_shifts(bits):
    """Chosen form on the Stockham kernel (run decode + MFMA first pass) with
    the quadratic delay model sweeping the integer shift past -N and +N:
    odd shifts, runs crossing the frame edges, frames shifted wholly out."""
    n, fs = 4096, 256e6
    c = make_case(n=n, fs=fs, nf=400, bits=bits, lsb=(True, False), no_table=True,
                  net_d1_base=-20.0e-6, total_rate1_base=6.5e-3, total_accel1_base=2e-3,
                  lo1_hz=3e9, seed=9 + bits[0], rho=0.4)
    d = R.make_delay_cfg(None, **c["delay"])
    ints = [R.frame_delay(i, d).int1 for i in range(c["nf"])]
    assert min(ints) < -n and max(ints) > n and any(i % 2 for i in ints)
    p = gpu_plan(c, F.PRECISION_CHOSEN)
    assert p.info.fft_form == 0 and p.info.fft_mfma == 1
    _compare(gpu_run(c, plan=p), oracle_run(c, R.MODE_F64), CHOSEN_TOL)


def test_device_entry_matches_host_entry_and_is_deterministic():
    c = make_case(n=1024, fs=64e6, nf=3000, seed=8, delay_user_samples=4.2, lo1_hz=1e9)
    p = gpu_plan(c)
    cfg = F.make_delay_cfg(c["table"], **c["delay"])
    a = p.sector_run(c["raw1"], c["raw2"], 0, c["nf"], cfg)
    b = p.sector_run(c["raw1"], c["raw2"], 0, c["nf"], cfg)
    for k in ("acc11", "acc22", "acc12"):
        assert np.array_equal(a[k], b[k])
    d1 = F.DeviceBuffer(c["raw1"].nbytes)
    d2 = F.DeviceBuffer(c["raw2"].nbytes)
    dacc = F.DeviceBuffer(F.N_PRODUCTS * p.half * 8)
    d1.upload(c["raw1"])
    d2.upload(c["raw2"])
    p.sector_run_device(d1.ptr, d2.ptr, 0, c["nf"], cfg, dacc.ptr)
    host = dacc.download(np.zeros(F.N_PRODUCTS * p.half))
    h = p.half
    assert np.array_equal(host[h:2 * h], a["acc11"])
    assert np.array_equal(host[2 * h:3 * h], a["acc22"])
    assert np.array_equal(host[3 * h:4 * h] + 1j * host[4 * h:], a["acc12"])
    p.close()


def test_timing_counts_launches():
    c = make_case(n=1024, fs=64e6, nf=500, seed=1)
    p = gpu_plan(c)
    p.set_timing(True)
    cfg = F.make_delay_cfg(c["table"], **c["delay"])
    for _ in range(3):
        p.sector_run(c["raw1"], c["raw2"], 0, c["nf"], cfg)
    ms, n = p.get_timing()
    assert n == 3 and ms > 0
    p.close()


# ------------------------------------------------- chosen form: FP16-MFMA FFT
def test_mfma_eligibility():
    lv2 = [LEVELS_2BIT, LEVELS_2BIT]
    assert F.Plan(1024, levels=lv2, df_hz=1.0, precision=F.PRECISION_CHOSEN).info.fft_form == 1
    assert F.Plan(1024, levels=lv2, df_hz=1.0).info.fft_form == 0          # plain: FP32 FFT
    assert F.Plan(2048, levels=lv2, df_hz=1.0, precision=F.PRECISION_CHOSEN).info.fft_form == 1
    assert F.Plan(4096, levels=lv2, df_hz=1.0, precision=F.PRECISION_CHOSEN).info.fft_form == 0
    lv8 = [np.arange(256, dtype=np.float64) - 127.5] * 2
    assert F.Plan(1024, bits=(8, 8), levels=lv8, df_hz=1.0, precision=F.PRECISION_CHOSEN).info.fft_form == 1
    # a 2048-sample 8-bit frame is 512 words, past the staging of one frame
    assert F.Plan(2048, bits=(8, 8), levels=lv8, df_hz=1.0, precision=F.PRECISION_CHOSEN).info.fft_form == 0
    p2 = F.Plan(2048, levels=lv2, df_hz=1.0, precision=F.PRECISION_CHOSEN)
    with pytest.raises(F.FxcError):
        F.plan_set_tuning(p2, variant=0)            # no whole-frame organization at N = 2048
    odd = [np.array([-1.1, -0.3, 0.3, 1.1])] * 2                           # not exact in FP16
    assert F.Plan(1024, levels=odd, df_hz=1.0, precision=F.PRECISION_CHOSEN).info.fft_form == 0
    # fft_mfma: the first DFT-32 on FP16 MFMA, also in the HBM four-step (N >= 32768)
    assert F.Plan(1024, levels=lv2, df_hz=1.0, precision=F.PRECISION_CHOSEN).info.fft_mfma == 1
    mid = F.Plan(4096, levels=lv2, df_hz=1.0, precision=F.PRECISION_CHOSEN).info
    assert mid.fft_form == 0 and mid.fft_mfma == 1                      # Stockham, MFMA first pass
    for n in (256, 512):                                                  # packed frames per slot
        small = F.Plan(n, levels=lv2, df_hz=1.0, precision=F.PRECISION_CHOSEN).info
        assert small.fft_form == 1 and small.fft_mfma == 1
    assert F.Plan(512, levels=lv2, df_hz=1.0).info.fft_mfma == 0
    assert F.Plan(128, levels=lv2, df_hz=1.0, precision=F.PRECISION_CHOSEN).info.fft_mfma == 0
    assert F.Plan(4096, levels=lv2, df_hz=1.0).info.fft_mfma == 0
    big = F.Plan(65536, levels=lv2, df_hz=1.0, precision=F.PRECISION_CHOSEN).info
    assert big.fft_form == 2 and big.fft_mfma == 1
    assert F.Plan(65536, levels=lv2, df_hz=1.0).info.fft_mfma == 0
    assert F.Plan(32768, levels=odd, df_hz=1.0, precision=F.PRECISION_CHOSEN).info.fft_mfma == 0


@pytest.mark.parametrize("variant", [0, 1])
def test_mfma_rfft_vs_numpy(variant):
    """Stage 2 as an FP32 MFMA GEMM (variant 0) or as the column DFT-16 in
    registers after a transposed LDS store (variant 1, the default).  The
    RMS bar also pins the FP16 hi + lo split of the DFT-32 coefficients: a
    lost lo part (e.g. FP16 subnormals flushed) gives errors near 1e-4."""
    n = 1024
    p = F.Plan(n, levels=[LEVELS_2BIT, LEVELS_2BIT], df_hz=1.0, precision=F.PRECISION_CHOSEN)
    assert p.info.fft_form == 1 and p.info.variant == 1
    F.plan_set_tuning(p, variant=variant)
    rng = np.random.default_rng(77)
    # FP16-exact inputs (2-bit levels), as the decode feeds the first stage
    x = LEVELS_2BIT[rng.integers(0, 4, size=(7, n))].astype(np.float32)
    x[0] = 0.0
    x[1] = 0.0
    x[1, 5] = 3.0                         # delayed impulse: exact linear phase
    got = p.rfft_frames(x)
    ref = np.fft.rfft(x.astype(np.float64), axis=1)
    scale = np.sqrt(n * 5.0)
    err = np.sqrt(np.mean(np.abs(got - ref) ** 2, axis=1)) / scale
    assert np.all(err[2:] < 3e-7 * np.sqrt(np.log2(n))), err
    assert np.max(np.abs(got[0])) == 0.0
    assert np.max(np.abs(got[1] - ref[1])) < 3e-6
    assert np.all(got[:, 0].imag == 0) and np.all(got[:, -1].imag == 0)


@pytest.mark.parametrize("variant", [0, 1, 2])
@pytest.mark.parametrize("bits", [(4, 1), (1, 8), (2, 4)])
def test_mfma_sector_mixed_bits(bits, variant):
    n, fs = 1024, 512e6
    rng = np.random.default_rng(sum(bits))
    c = make_case(n=n, fs=fs, nf=3000, bits=bits, shuffles=(rng.permutation(32), None),
                  lsb=(True, False), rho=0.5, delay_samples=-3.4, seed=21,
                  delay_user_samples=-3.4, extra_delay_rate_sps=2e-9, lo1_hz=8.4e9,
                  lo2_hz=8.4e9)
    p = gpu_plan(c, F.PRECISION_CHOSEN)
    assert p.info.fft_form == 1
    F.plan_set_tuning(p, variant=variant)
    _compare(gpu_run(c, plan=p), oracle_run(c, R.MODE_F64), CHOSEN_TOL)


@pytest.mark.parametrize("variant", [0, 1, 2])
@pytest.mark.parametrize("grid", [0, 1])
@pytest.mark.parametrize("overlap_only", [False, True])
def test_mfma_partial_band_grids(grid, overlap_only, variant):
    """Common band start with a partial overlap (bins 100..399): the grid
    station's ACF covers every bin, the other station's and the XCF only
    the overlap (or both ACFs only the overlap)."""
    n, fs = 1024, 256e6
    prods = F.PRODUCT_ACF | F.PRODUCT_XCF
    if overlap_only:
        prods |= F.PRODUCT_ACF_OVERLAP_ONLY
    c = make_case(n=n, fs=fs, nf=2000, band=(0, 100, 400, 100), grid=grid, products=prods,
                  rho=0.5, delay_samples=2.6, seed=31, delay_user_samples=2.6, lo1_hz=8e9,
                  lo2_hz=8e9)
    p = gpu_plan(c, F.PRECISION_CHOSEN)
    assert p.info.fft_form == 1
    F.plan_set_tuning(p, variant=variant)
    got = gpu_run(c, plan=p)
    ref = oracle_run(c, R.MODE_F64)
    _compare(got, ref, CHOSEN_TOL)
    for k in ("acc11", "acc22", "acc12"):
        assert zero_pattern_equal(got[k], ref[k]), k


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_mfma_large_and_odd_shifts(variant):
    """Analytic delay model sweeping the integer shift from about -1100 to
    +200 samples: odd shifts, shifts past the staging pad and whole frames
    shifted out."""
    n, fs = 1024, 128e6
    c = make_case(n=n, fs=fs, nf=1500, no_table=True, net_d1_base=-8.6e-6,
                  total_rate1_base=8.5e-4, lo1_hz=2e9, seed=6, rho=0.3)
    d = R.make_delay_cfg(None, **c["delay"])
    ints = [R.frame_delay(i, d).int1 for i in range(c["nf"])]
    assert min(ints) < -n and max(ints) > 64 and any(i % 2 for i in ints)
    p = gpu_plan(c, F.PRECISION_CHOSEN)
    assert p.info.fft_form == 1
    F.plan_set_tuning(p, variant=variant)
    _compare(gpu_run(c, plan=p), oracle_run(c, R.MODE_F64), CHOSEN_TOL)


@pytest.mark.parametrize("nf", [1, 3, 5, 37, 2051])
def test_mfma_kernel_organizations_agree(nf):
    """Whole-frame waves (FP32 MFMA stage 2), producer/consumer waves (VALU
    radix-4 x 4 stage 2) and the same with FP32 batch sums in the consumers
    agree to FP32 rounding for frame counts that
    leave producers (or whole workgroups) without frames, and both match
    the FP64 oracle.  With a single frame a bin of tiny power turns FP32
    rounding into a per-bin relative difference near 1e-5, hence the
    chosen-form bar for the agreement too."""
    n, fs = 1024, 1024e6
    rng = np.random.default_rng(nf)
    c = make_case(n=n, fs=fs, nf=nf, shuffles=(None, rng.permutation(32)), lsb=(False, True),
                  rho=0.4, delay_samples=1.3, seed=40 + nf, delay_user_samples=1.3,
                  extra_delay_rate_sps=1e-9, lo1_hz=6.7e9, lo2_hz=6.7e9)
    res = []
    for v in (0, 1, 2):
        p = gpu_plan(c, F.PRECISION_CHOSEN)
        assert p.info.fft_form == 1
        F.plan_set_tuning(p, variant=v)
        res.append(gpu_run(c, plan=p))
    for r in res[1:]:
        e = spectra_err(r, res[0])
        assert max(e.values()) < CHOSEN_TOL, e
    o64 = oracle_run(c, R.MODE_F64)
    for r in res:
        _compare(r, o64, CHOSEN_TOL)


# N = 2048: 32 x 32 four-step (one station-frame per producer slot, column
# DFT-32 split over lane pairs); producer/consumer organization only.
@pytest.mark.parametrize("bits", [(4, 1), (1, 2), (2, 4), (2, 2)])
def test_mfma2048_sector_mixed_bits(bits):
    n, fs = 2048, 512e6
    rng = np.random.default_rng(100 + sum(bits))
    c = make_case(n=n, fs=fs, nf=1700, bits=bits, shuffles=(rng.permutation(32), None),
                  lsb=(True, False), rho=0.5, delay_samples=-3.4, seed=23,
                  delay_user_samples=-3.4, extra_delay_rate_sps=2e-9, lo1_hz=8.4e9,
                  lo2_hz=8.4e9)
    p = gpu_plan(c, F.PRECISION_CHOSEN)
    assert p.info.fft_form == 1 and p.info.variant == 1
    _compare(gpu_run(c, plan=p), oracle_run(c, R.MODE_F64), CHOSEN_TOL)


@pytest.mark.parametrize("grid", [0, 1])
@pytest.mark.parametrize("overlap_only", [False, True])
def test_mfma2048_partial_band_grids(grid, overlap_only):
    n, fs = 2048, 256e6
    prods = F.PRODUCT_ACF | F.PRODUCT_XCF
    if overlap_only:
        prods |= F.PRODUCT_ACF_OVERLAP_ONLY
    c = make_case(n=n, fs=fs, nf=1200, band=(0, 200, 700, 200), grid=grid, products=prods,
                  rho=0.5, delay_samples=2.6, seed=33, delay_user_samples=2.6, lo1_hz=8e9,
                  lo2_hz=8e9)
    p = gpu_plan(c, F.PRECISION_CHOSEN)
    assert p.info.fft_form == 1
    got = gpu_run(c, plan=p)
    ref = oracle_run(c, R.MODE_F64)
    _compare(got, ref, CHOSEN_TOL)
    for k in ("acc11", "acc22", "acc12"):
        assert zero_pattern_equal(got[k], ref[k]), k


def test_mfma2048_large_and_odd_shifts():
    n, fs = 2048, 128e6
    c = make_case(n=n, fs=fs, nf=1000, no_table=True, net_d1_base=-17.2e-6,
                  total_rate1_base=1.2e-3, lo1_hz=2e9, seed=7, rho=0.3)
    d = R.make_delay_cfg(None, **c["delay"])
    ints = [R.frame_delay(i, d).int1 for i in range(c["nf"])]
    assert min(ints) < -n and max(ints) > 64 and any(i % 2 for i in ints)
    p = gpu_plan(c, F.PRECISION_CHOSEN)
    assert p.info.fft_form == 1
    _compare(gpu_run(c, plan=p), oracle_run(c, R.MODE_F64), CHOSEN_TOL)


@pytest.mark.parametrize("nf", [1, 3, 5, 37, 2051])
def test_mfma2048_ragged_frame_counts(nf):
    """Frame counts that leave producer slots, streams or whole workgroups
    without frames; the MFMA kernel against the FP32 Stockham kernel (the
    plain form) and both against the FP64 oracle."""
    n, fs = 2048, 1024e6
    rng = np.random.default_rng(nf)
    c = make_case(n=n, fs=fs, nf=nf, shuffles=(None, rng.permutation(32)), lsb=(False, True),
                  rho=0.4, delay_samples=1.3, seed=50 + nf, delay_user_samples=1.3,
                  extra_delay_rate_sps=1e-9, lo1_hz=6.7e9, lo2_hz=6.7e9)
    p = gpu_plan(c, F.PRECISION_CHOSEN)
    assert p.info.fft_form == 1
    got = gpu_run(c, plan=p)
    o64 = oracle_run(c, R.MODE_F64)
    _compare(got, o64, CHOSEN_TOL)
    _check_plain(c, gpu_run(c, plan=gpu_plan(c, F.PRECISION_PLAIN_FP32)))


# N = 256 / 512: four (two) station-frames packed into one producer slot,
# their 32-point rows filling one DFT-32 tile; producer/consumer
# organization only.
@pytest.mark.parametrize("n", [256, 512])
@pytest.mark.parametrize("bits", [(4, 1), (1, 2), (8, 2), (2, 2)])
def test_mfma_packed_sector_mixed_bits(n, bits):
    fs = 512e6
    rng = np.random.default_rng(200 + n + sum(bits))
    c = make_case(n=n, fs=fs, nf=3000, bits=bits, shuffles=(rng.permutation(32), None),
                  lsb=(True, False), rho=0.5, delay_samples=-3.4, seed=29,
                  delay_user_samples=-3.4, extra_delay_rate_sps=2e-9, lo1_hz=8.4e9,
                  lo2_hz=8.4e9)
    p = gpu_plan(c, F.PRECISION_CHOSEN)
    assert p.info.fft_form == 1 and p.info.variant == 1
    _compare(gpu_run(c, plan=p), oracle_run(c, R.MODE_F64), CHOSEN_TOL)


@pytest.mark.parametrize("n", [256, 512])
@pytest.mark.parametrize("grid", [0, 1])
@pytest.mark.parametrize("overlap_only", [False, True])
def test_mfma_packed_partial_band_grids(n, grid, overlap_only):
    fs = 256e6
    prods = F.PRODUCT_ACF | F.PRODUCT_XCF
    if overlap_only:
        prods |= F.PRODUCT_ACF_OVERLAP_ONLY
    band = (0, 20, n // 2 - 30, 20)
    c = make_case(n=n, fs=fs, nf=2000, band=band, grid=grid, products=prods, rho=0.5,
                  delay_samples=2.6, seed=37, delay_user_samples=2.6, lo1_hz=8e9, lo2_hz=8e9)
    p = gpu_plan(c, F.PRECISION_CHOSEN)
    assert p.info.fft_form == 1
    got = gpu_run(c, plan=p)
    ref = oracle_run(c, R.MODE_F64)
    _compare(got, ref, CHOSEN_TOL)
    for k in ("acc11", "acc22", "acc12"):
        assert zero_pattern_equal(got[k], ref[k]), k


@pytest.mark.parametrize("n", [256, 512])
def test_mfma_packed_large_and_odd_shifts(n):
    """Integer shifts beyond the staged frames' padding (the checked
    operand path), beyond +-N (all-zero frames) and of both parities."""
    fs = 128e6
    base, rate = (-2.8e-6, 1.3e-3) if n == 256 else (-5e-6, 1.17e-3)
    c = make_case(n=n, fs=fs, nf=1500, no_table=True, net_d1_base=base, total_rate1_base=rate,
                  lo1_hz=2e9, seed=9, rho=0.3)
    d = R.make_delay_cfg(None, **c["delay"])
    ints = [R.frame_delay(i, d).int1 for i in range(c["nf"])]
    assert min(ints) < -n and max(ints) > 40 and any(i % 2 for i in ints)
    assert any(-32 <= i <= 32 for i in ints)
    p = gpu_plan(c, F.PRECISION_CHOSEN)
    assert p.info.fft_form == 1
    _compare(gpu_run(c, plan=p), oracle_run(c, R.MODE_F64), CHOSEN_TOL)


@pytest.mark.parametrize("n", [256, 512])
@pytest.mark.parametrize("nf", [1, 3, 5, 37, 2051])
def test_mfma_packed_ragged_frame_counts(n, nf):
    """Frame counts that leave slots partly filled and producers, streams or
    workgroups without frames; against the plain form and the FP64 oracle."""
    fs = 1024e6
    rng = np.random.default_rng(nf + n)
    c = make_case(n=n, fs=fs, nf=nf, shuffles=(None, rng.permutation(32)), lsb=(False, True),
                  rho=0.4, delay_samples=1.3, seed=60 + nf, delay_user_samples=1.3,
                  extra_delay_rate_sps=1e-9, lo1_hz=6.7e9, lo2_hz=6.7e9)
    p = gpu_plan(c, F.PRECISION_CHOSEN)
    assert p.info.fft_form == 1
    _compare(gpu_run(c, plan=p), oracle_run(c, R.MODE_F64), CHOSEN_TOL)
    _check_plain(c, gpu_run(c, plan=gpu_plan(c, F.PRECISION_PLAIN_FP32)))


@pytest.mark.parametrize("n", [256, 512])
def test_mfma_packed_many_batches(n):
    """Device-generated stations long enough for several 64-frame batches per
    producer (and slots that straddle them), against the FP64 oracle."""
    fs, bits = 1024e6, (2, 4)
    nf = 150_000 if n == 256 else 80_000
    sh1 = np.random.default_rng(n).permutation(32).astype(np.int32)
    s = F.Stream()
    d1 = F.DeviceBuffer(nf * n * bits[0] // 8)
    d2 = F.DeviceBuffer(nf * n * bits[1] // 8)
    F.synth_station_device(d1.ptr, nf * n, bits[0], sh1, False, 91, 0, 0.4, 3, s.ptr)
    F.synth_station_device(d2.ptr, nf * n, bits[1], None, True, 91, 1, 0.4, 3, s.ptr)
    raw1 = d1.download(np.zeros(nf * n * bits[0] // 8, dtype=np.uint8), s.ptr)
    raw2 = d2.download(np.zeros(nf * n * bits[1] // 8, dtype=np.uint8), s.ptr)
    s.sync()
    lv = [np.arange(1 << b, dtype=np.float64) - ((1 << b) - 1) / 2.0 for b in bits]
    table = np.zeros(int(np.ceil(nf * n / fs)) + 2)
    delay = dict(frame_dt=n / fs, fs=fs, delay_user_samples=2.7, extra_delay_rate_sps=3e-9,
                 lo1_hz=5e9, lo2_hz=5e9)
    c = dict(n=n, fs=fs, nf=nf, bits=bits, levels=lv, shuffles=(sh1, IDENT), lsb=(False, True),
             band=(0, 0, n // 2, 0), grid=0, products=F.PRODUCT_ACF | F.PRODUCT_XCF, raw1=raw1,
             raw2=raw2, delay=delay, table=table, first_frame=0)
    p = gpu_plan(c, F.PRECISION_CHOSEN)
    assert p.info.fft_form == 1
    _compare(gpu_run(c, plan=p), oracle_run(c, R.MODE_F64), CHOSEN_TOL)
    d1.free()
    d2.free()


def test_mfma_needs_common_band_start():
    lv2 = [LEVELS_2BIT, LEVELS_2BIT]
    p = F.Plan(1024, levels=lv2, df_hz=1.0, band=(-16, 16, 512, 0),
               precision=F.PRECISION_CHOSEN)
    assert p.info.fft_form == 0 and p.info.fft_mfma == 1   # Stockham with the MFMA first pass


@pytest.mark.parametrize("n,variant", [(1024, 0), (2048, 0), (2048, 1), (4096, 0), (4096, 1),
                                       (8192, 1), (16384, 0)])
def test_rfft_stockham_mfma_vs_numpy(n, variant):
    """Chosen form on the Stockham kernel: the first pass (radix 32) as FP16
    hi + lo DFT-32 MFMA GEMMs, the rest in FP32.  FP16-exact inputs as the
    decode feeds it; the RMS bar pins the lo part."""
    lv2 = [LEVELS_2BIT, LEVELS_2BIT]
    band = (-16, 16, n // 2, 0) if n <= 2048 else None   # keeps N = 1024/2048 off the fused kernels
    kw = dict(band=band) if band else {}
    p = F.Plan(n, levels=lv2, df_hz=1.0, precision=F.PRECISION_CHOSEN, **kw)
    assert p.info.fft_form == 0 and p.info.fft_mfma == 1
    if variant:
        F.plan_set_tuning(p, variant=variant)
    rng = np.random.default_rng(n + variant)
    x = LEVELS_2BIT[rng.integers(0, 4, size=(5, n))].astype(np.float32)
    x[0] = 0.0
    x[0, 5] = 3.0                                      # impulse: exact linear phase
    got = p.rfft_frames(x)
    ref = np.fft.rfft(x.astype(np.float64), axis=1)
    err = np.sqrt(np.mean(np.abs(got - ref) ** 2, axis=1) / np.mean(np.abs(ref) ** 2, axis=1))
    assert np.all(err < 3e-7 * np.log2(n)), err
    assert np.max(np.abs(got[0] - ref[0])) < 3e-6
    assert np.all(got[:, 0].imag == 0) and np.all(got[:, -1].imag == 0)


# ------------------------------------------------------------ phased array
def _phased_desc(c, w, bit_out, output_lsb):
    n = c["n"]
    return dict(fft_len=n, bits1=c["bits"][0], bits2=c["bits"][1], levels1=c["levels"][0],
                levels2=c["levels"][1], shuffle1=c["shuffles"][0], shuffle2=c["shuffles"][1],
                lsb1=int(c["lsb"][0]), lsb2=int(c["lsb"][1]), a1s=c["band"][1], a1e=c["band"][2],
                a2s=c["band"][3], out_grid=c["grid"], df_hz=c["fs"] / n, need_acf=0, need_xcf=0,
                need_phased=1, write_raw=1, w1=w[0], w2=w[1], bit_out=bit_out,
                output_lsb=int(output_lsb), cpu_threads=8)


PHASED_CASES = [
    (256, (2, 8), 1, 0, True),         # chosen: the packed-frame MFMA kernel
    (512, (4, 2), 0, 0, False),
    (1024, (2, 2), 0, 0, False),
    (1024, (8, 1), 1, 0, True),        # chosen: the producer/consumer MFMA kernel, other grid
    (1024, (4, 4), 0, 8, False),       # band shift: the chosen form on the Stockham kernel
    (1024, (3, 2), 1, 0, True),        # 3-bit output: codes straddle the output words
    (512, (4, 2), 1, 8, True),
    (2048, (2, 1), 0, -16, False),
    (4096, (8, 2), 0, 64, True),
    (16384, (1, 8), 1, -32, False),
    (32768, (4, 1), 0, 2048, True),
    (65536, (2, 2), 1, -64, False),
]


def _phased_case(n, bits, grid, band_shift, lsb_out, precision):
    fs = 128e6
    half = n // 2                                  # main.rs:360-372 with v1 = v2 = N/2
    sb = band_shift
    a1s = max(0, -sb)
    band = (sb, a1s, min(half, half - sb), a1s + sb)
    rng = np.random.default_rng(n + bits[0])
    nf = 1500 if n <= 16384 else 48               # N >= 32768: the HBM four-step path
    c = make_case(n=n, fs=fs, nf=nf, bits=bits, shuffles=(rng.permutation(32), None),
                  lsb=(False, True), band=band, grid=grid, rho=0.4, delay_samples=3.3, seed=41,
                  delay_user_samples=3.3, extra_delay_rate_sps=1e-9, lo1_hz=5e9, lo2_hz=5e9,
                  products=F.PRODUCT_PHASED | F.PRODUCT_RAW_OUT)
    w = (0.7, 1.3)
    bit_out = max(bits)
    p = F.Plan(n, bits=bits, levels=c["levels"], shuffles=c["shuffles"], lsb=c["lsb"],
               band=band, df_hz=fs / n, products=F.PRODUCT_PHASED | F.PRODUCT_RAW_OUT,
               out_grid=grid, w=w, bit_out=bit_out, output_lsb=lsb_out, precision=precision)
    cfg = F.make_delay_cfg(c["table"], **c["delay"])
    got = p.sector_run(c["raw1"], c["raw2"], 0, c["nf"], cfg)
    mode, tol = ((R.MODE_F32, PLAIN_TOL) if precision == F.PRECISION_PLAIN_FP32
                 else (R.MODE_F64, CHOSEN_TOL))
    ref = R.sector_run(_phased_desc(c, w, bit_out, lsb_out), c["raw1"], c["raw2"], 0, c["nf"],
                       R.make_delay_cfg(c["table"], **c["delay"]), mode=mode, threads=8)
    assert got["n_failed"] == 0 and ref["n_failed"] == 0
    e = rel_err(got["acc_ph"], ref["acc_ph"])
    print(f"[phased] N={n} precision {precision}: acc_ph {e:.2e}", flush=True)
    assert e < tol
    for k in ("acc11", "acc22"):
        assert np.all(got[k] == 0)
    assert np.all(got["acc12"] == 0)
    bpf_o = n * bit_out // 8
    assert got["raw_out"].size == c["nf"] * bpf_o
    dp = out_decode_plan(bit_out, c["shuffles"][0], c["levels"][0])
    diff = 0
    for f in range(0, c["nf"], 7):
        seg = slice(f * bpf_o, (f + 1) * bpf_o)
        _, vg = dp.decode(got["raw_out"][seg], n, False)
        _, vr = dp.decode(ref["raw_out"][seg], n, False)
        diff += int(np.count_nonzero(vg != vr))
    assert diff <= 1e-4 * n * len(range(0, c["nf"], 7)) + 2, diff


@pytest.mark.parametrize("n,bits,grid,band_shift,lsb_out", PHASED_CASES)
def test_phased_vs_oracle(n, bits, grid, band_shift, lsb_out):
    """yi-phasedarray sector: |w1 s1 + w2 s2|^2 within the plain bar of the
    oracle's phased branch at every N (each station's rotation is the
    reference's own f32 chain over all N/2+1 bins); the .raw output decodes
    to the same levels except at rare quantiser decision boundaries."""
    _phased_case(n, bits, grid, band_shift, lsb_out, F.PRECISION_PLAIN_FP32)


@pytest.mark.parametrize("n,bits,grid,band_shift,lsb_out", PHASED_CASES)
def test_phased_chosen_vs_fp64_oracle(n, bits, grid, band_shift, lsb_out):
    """The chosen phased form (rotations and fringe phasors from FP64
    phases, no f32 chain table) against the oracle's FP64 phased branch:
    |cb|^2 within the chosen bar, .raw codes equal except at rare quantiser
    decision boundaries."""
    _phased_case(n, bits, grid, band_shift, lsb_out, F.PRECISION_CHOSEN)


@pytest.mark.parametrize("n,nf,bits,grid,band,lsb_out", [
    (1024, 3000, (4, 2), 1, (0, 16, 400, 16), True),        # partial overlap on a common band start
    (1024, 140_000, (2, 1), 0, (0, 0, 512, 0), False),      # two .raw chunks, 64-frame batches
    (256, 5001, (8, 1), 0, (0, 9, 100, 9), True),           # packed frames, partial slots
    (256, 530_000, (1, 2), 1, (0, 0, 128, 0), False),       # packed frames, two .raw chunks
    (512, 270_000, (2, 4), 0, (0, 0, 256, 0), True),
])
def test_phased_chosen_mfma_kernel(n, nf, bits, grid, band, lsb_out):
    """Chosen phased sum at N = 256 / 512 / 1024 on the producer/consumer
    MFMA kernels (both stations on a common band start).  Device-generated
    stations; |cb|^2 within the chosen bar of the oracle's FP64 phased branch
    over every frame, and the .raw codes of sampled frames (every frame
    around the first chunk boundary) equal except at rare quantiser decision
    boundaries."""
    fs = 256e6
    sh1 = np.random.default_rng(nf).permutation(32).astype(np.int32)
    s = F.Stream()
    d1 = F.DeviceBuffer(nf * n * bits[0] // 8)
    d2 = F.DeviceBuffer(nf * n * bits[1] // 8)
    F.synth_station_device(d1.ptr, nf * n, bits[0], sh1, True, 77, 0, 0.4, 3, s.ptr)
    F.synth_station_device(d2.ptr, nf * n, bits[1], None, False, 77, 1, 0.4, 3, s.ptr)
    raw1 = d1.download(np.zeros(nf * n * bits[0] // 8, dtype=np.uint8), s.ptr)
    raw2 = d2.download(np.zeros(nf * n * bits[1] // 8, dtype=np.uint8), s.ptr)
    s.sync()
    lv = [np.arange(1 << b, dtype=np.float64) - ((1 << b) - 1) / 2.0 for b in bits]
    ident = np.arange(32, dtype=np.int32)
    table = np.zeros(int(np.ceil(nf * n / fs)) + 2)
    delay = dict(frame_dt=n / fs, fs=fs, delay_user_samples=-2.2, extra_delay_rate_sps=2e-9,
                 lo1_hz=3e9, lo2_hz=3e9)
    w = (0.9, 1.2)
    bit_out = max(bits)
    p = F.Plan(n, bits=bits, levels=lv, shuffles=(sh1, None), lsb=(True, False), band=band,
               df_hz=fs / n, products=F.PRODUCT_PHASED | F.PRODUCT_RAW_OUT, out_grid=grid, w=w,
               bit_out=bit_out, output_lsb=lsb_out, precision=F.PRECISION_CHOSEN)
    assert p.info.fft_form == 1
    acc = F.DeviceBuffer(F.N_PRODUCTS * p.half * 8)
    rawo = F.DeviceBuffer(nf * p.info.bpf_out)
    p.sector_run_device(d1.ptr, d2.ptr, 0, nf, F.make_delay_cfg(table, **delay), acc.ptr, rawo.ptr,
                        stream=s.ptr)
    got_acc = acc.download(np.zeros(F.N_PRODUCTS * p.half), s.ptr)
    got_raw = rawo.download(np.zeros(nf * p.info.bpf_out, dtype=np.uint8), s.ptr)
    s.sync()
    desc = dict(fft_len=n, bits1=bits[0], bits2=bits[1], levels1=lv[0], levels2=lv[1],
                shuffle1=sh1, shuffle2=ident, lsb1=1, lsb2=0, a1s=band[1], a1e=band[2],
                a2s=band[3], out_grid=grid, df_hz=fs / n, need_acf=0, need_xcf=0, need_phased=1,
                write_raw=1, w1=w[0], w2=w[1], bit_out=bit_out, output_lsb=int(lsb_out),
                cpu_threads=8)
    ref = R.sector_run(desc, raw1, raw2, 0, nf, R.make_delay_cfg(table, **delay), mode=R.MODE_F64,
                       threads=8)
    h = p.half
    e = rel_err(got_acc[:h], ref["acc_ph"])
    print(f"[phased-mf] nf={nf} bits={bits}: acc_ph {e:.2e}", flush=True)
    assert e < CHOSEN_TOL
    assert np.all(got_acc[h:] == 0)
    bpf_o = n * bit_out // 8
    cmax = (512 << 20) // ((n // 2 + 1) * 8)           # most frames per .raw chunk
    nchunks = -(-nf // cmax)
    edge = -(-nf // nchunks)                           # equal chunks: the first boundary
    frames = sorted(set(range(0, nf, max(1, nf // 200))) |
                    set(range(max(0, edge - 40), min(nf, edge + 40))))
    dp = out_decode_plan(bit_out, sh1, lv[0])
    diff = 0
    for f in frames:
        seg = slice(f * bpf_o, (f + 1) * bpf_o)
        _, vg = dp.decode(got_raw[seg], n, False)
        _, vr = dp.decode(ref["raw_out"][seg], n, False)
        diff += int(np.count_nonzero(vg != vr))
    assert diff <= 1e-4 * n * len(frames) + 2, diff
    for b in (d1, d2, acc, rawo):
        b.free()


@pytest.mark.parametrize("nf,chunk", [(5, 0), (23, 7)])
def test_phased_large_chunks_and_few_frames(nf, chunk):
    """N = 32768 phased sum with fewer frames than frame groups, and over
    several scratch chunks (partials accumulated in order, rotation chains
    computed for whole chunks), against the oracle's phased branch."""
    n, fs, bits = 32768, 256e6, (2, 4)
    band = (16, 0, n // 2 - 16, 16)
    c = make_case(n=n, fs=fs, nf=nf, bits=bits, lsb=(True, False), band=band, rho=0.4,
                  delay_samples=-2.7, seed=61 + nf, delay_user_samples=-2.7, lo1_hz=4e9,
                  lo2_hz=4e9, products=F.PRODUCT_PHASED | F.PRODUCT_RAW_OUT)
    w = (1.1, 0.8)
    p = F.Plan(n, bits=bits, levels=c["levels"], shuffles=c["shuffles"], lsb=c["lsb"], band=band,
               df_hz=fs / n, products=F.PRODUCT_PHASED | F.PRODUCT_RAW_OUT, w=w, bit_out=4)
    if chunk:
        F.plan_set_tuning(p, split_m1=128, chunk_frames=chunk)
    cfg = F.make_delay_cfg(c["table"], **c["delay"])
    got = p.sector_run(c["raw1"], c["raw2"], 0, nf, cfg)
    ref = R.sector_run(_phased_desc(c, w, 4, False), c["raw1"], c["raw2"], 0, nf,
                       R.make_delay_cfg(c["table"], **c["delay"]), mode=R.MODE_F32, threads=8)
    assert rel_err(got["acc_ph"], ref["acc_ph"]) < PLAIN_TOL
    dp = out_decode_plan(4, c["shuffles"][0], c["levels"][0])
    bpf_o = n * 4 // 8
    diff = 0
    for f in range(nf):
        seg = slice(f * bpf_o, (f + 1) * bpf_o)
        _, vg = dp.decode(got["raw_out"][seg], n, False)
        _, vr = dp.decode(ref["raw_out"][seg], n, False)
        diff += int(np.count_nonzero(vg != vr))
    assert diff <= 1e-4 * n * nf + 2, diff


@pytest.mark.parametrize("n,nf", [(1024, 700), (32768, 40)])
def test_phased_no_raw_out_matches_with_raw_out(n, nf):
    fs = 64e6
    c = make_case(n=n, fs=fs, nf=nf, rho=0.3, seed=3, lo1_hz=2e9)
    cfg = F.make_delay_cfg(c["table"], **c["delay"])
    res = []
    for prods in (F.PRODUCT_PHASED, F.PRODUCT_PHASED | F.PRODUCT_RAW_OUT):
        p = F.Plan(n, levels=c["levels"], df_hz=fs / n, products=prods, bit_out=2)
        res.append(p.sector_run(c["raw1"], c["raw2"], 0, c["nf"], cfg)["acc_ph"])
    assert np.array_equal(res[0], res[1])


# --------------------------------------------- N = 32768 / 65536 (HBM four-step)
@pytest.mark.parametrize("n", [32768, 65536])
def test_rfft_large_vs_numpy(n):
    p = F.Plan(n, levels=[LEVELS_2BIT, LEVELS_2BIT], df_hz=1.0)
    assert p.info.fft_form == 2
    rng = np.random.default_rng(n)
    x = rng.standard_normal((3, n)).astype(np.float32)
    x[0] = 0.0
    x[0, 7] = 1.0                                      # impulse: exact linear phase
    got = p.rfft_frames(x)
    ref = np.fft.rfft(x.astype(np.float64), axis=1)
    err = np.sqrt(np.mean(np.abs(got - ref) ** 2, axis=1) / np.mean(np.abs(ref) ** 2, axis=1))
    assert np.all(err < 3e-7 * np.log2(n)), err
    assert np.all(got[:, 0].imag == 0) and np.all(got[:, -1].imag == 0)


@pytest.mark.parametrize("n,m1", [(32768, 128), (65536, 128), (65536, 256)])
def test_rfft_large_mfma_vs_numpy(n, m1):
    """Chosen form, N >= 32768: the first pass's column DFTs (M1 = 32 R) as
    an FP16 hi + lo DFT-32 MFMA GEMM, the W_M1 twiddles and DFT-R in FP32.
    FP16-exact inputs as the decode feeds it; the RMS bar pins the lo part
    (a lost lo part gives errors near 1e-4)."""
    p = F.Plan(n, levels=[LEVELS_2BIT, LEVELS_2BIT], df_hz=1.0, precision=F.PRECISION_CHOSEN)
    assert p.info.fft_form == 2 and p.info.fft_mfma == 1
    F.plan_set_tuning(p, split_m1=m1, chunk_frames=3)
    rng = np.random.default_rng(n + m1)
    x = LEVELS_2BIT[rng.integers(0, 4, size=(4, n))].astype(np.float32)
    x[0] = 0.0
    x[0, 7] = 3.0                                      # impulse: exact linear phase
    got = p.rfft_frames(x)
    ref = np.fft.rfft(x.astype(np.float64), axis=1)
    err = np.sqrt(np.mean(np.abs(got - ref) ** 2, axis=1) / np.mean(np.abs(ref) ** 2, axis=1))
    assert np.all(err < 3e-7 * np.log2(n)), err
    assert np.max(np.abs(got[0] - ref[0])) < 3e-6
    assert np.all(got[:, 0].imag == 0) and np.all(got[:, -1].imag == 0)


def test_512msps_4bit_1bit_fft65536_band_shift():
    """512 Msps, A 4-bit (16 uniform levels) / B 1-bit, FFT 65536,
    16 MHz centre offset -> band shift 2048 bins, overlap 30720 (main.rs:
    360-372), both output grids' bins; 40 frames of it."""
    n, fs = 65536, 512e6
    band = F.band_alignment(n, fs, 8400.0, 8384.0, 256.0, 256.0)
    assert band == (2048, 0, 30720, 2048)
    c = make_case(n=n, fs=fs, nf=40, bits=(4, 1), band=band, rho=0.3, delay_samples=1.7,
                  seed=13, delay_user_samples=1.7, extra_delay_rate_sps=1e-9, lo1_hz=8.4e9,
                  lo2_hz=8.384e9)
    o64 = oracle_run(c, R.MODE_F64)
    o32 = oracle_run(c, R.MODE_F32)
    g = gpu_run(c)
    # the plain form replays the reference's 30720-link f32 phase chain
    _compare(g, o32, PLAIN_TOL)
    # which itself walks away from FP64 truth by up to ~1e-3 over the band
    assert spectra_err(g, o64)["acc12"] < 1.2e-7 * 30720 + 1e-5
    gc = gpu_run(c, F.PRECISION_CHOSEN)
    _compare(gc, o64, CHOSEN_TOL)
    for k in ("acc11", "acc22", "acc12"):
        assert zero_pattern_equal(g[k], o64[k]), k


def test_large_chunking_is_a_sum():
    """A sector longer than one scratch chunk equals the sum of its halves
    run separately (partials accumulate over chunks in order)."""
    n, fs, nf = 32768, 256e6, 2600
    s = F.Stream()
    bpf = n * 2 // 8
    r1, r2 = F.DeviceBuffer(nf * bpf), F.DeviceBuffer(nf * bpf)
    F.synth_station_device(r1.ptr, nf * n, 2, None, False, 5, 0, 0.2, 3, s.ptr)
    F.synth_station_device(r2.ptr, nf * n, 2, None, False, 5, 1, 0.2, 3, s.ptr)
    s.sync()
    p = F.Plan(n, levels=[LEVELS_2BIT, LEVELS_2BIT], df_hz=fs / n,
               precision=F.PRECISION_CHOSEN)
    cfg = F.make_delay_cfg(np.zeros(4), frame_dt=n / fs, fs=fs, delay_user_samples=3.2,
                           lo1_hz=3e9)
    h = p.half
    acc = F.DeviceBuffer(3 * F.N_PRODUCTS * h * 8)
    p.sector_run_device(r1.ptr, r2.ptr, 0, nf, cfg, acc.ptr, stream=s.ptr)
    k = 1300
    p.sector_run_device(r1.ptr, r2.ptr, 0, k, cfg, acc.ptr + F.N_PRODUCTS * h * 8, stream=s.ptr)
    p.sector_run_device(r1.ptr + k * bpf, r2.ptr + k * bpf, k, nf - k, cfg,
                        acc.ptr + 2 * F.N_PRODUCTS * h * 8, stream=s.ptr)
    s.sync()
    out = acc.download(np.zeros(3 * F.N_PRODUCTS * h)).reshape(3, F.N_PRODUCTS, h)
    full, parts = out[0], out[1] + out[2]
    scale = np.abs(full).max(axis=1, keepdims=True)
    assert np.all(np.abs(full - parts) <= 1e-11 * scale + 1e-300)
    assert np.all(full[1] > 0)


# ---------------------------------------------------------------- autotuner
@pytest.mark.parametrize("n,variant,div", [(2048, 1, 1), (4096, 1, 2), (8192, 0, 4), (8192, 1, 1)])
def test_tuning_candidates_keep_parity(n, variant, div):
    """Every configuration the autotuner can pick gives the same spectra."""
    c = make_case(n=n, fs=128e6, nf=max(128, 2 ** 19 // n), rho=0.5, seed=n + variant,
                  delay_user_samples=2.3, lo1_hz=3e9)
    o64 = oracle_run(c, R.MODE_F64)
    for prec in (F.PRECISION_PLAIN_FP32, F.PRECISION_CHOSEN):
        p = gpu_plan(c, prec)
        F.plan_set_tuning(p, variant=variant, grid_divisor=div)
        g = gpu_run(c, plan=p)
        if prec == F.PRECISION_PLAIN_FP32:
            _check_plain(c, g)
        else:
            _compare(g, o64, CHOSEN_TOL)


def test_large_split_and_chunk_candidates():
    """Every large-form configuration the autotuner can pick: the fused
    four-step (variant 1, M1 = 128) and the four-step through HBM (variant
    0) at both splits, over one or several chunks."""
    n, fs = 65536, 512e6
    c = make_case(n=n, fs=fs, nf=40, bits=(4, 1), rho=0.3, seed=17, delay_user_samples=1.2,
                  lo1_hz=8e9)
    o64 = oracle_run(c, R.MODE_F64)
    p = gpu_plan(c, F.PRECISION_CHOSEN)
    assert p.info.variant == 1                        # the fused form is the default
    for v, m1, chunk in ((1, 128, 40), (1, 128, 7), (0, 256, 12), (0, 128, 40), (0, 128, 7)):
        p = gpu_plan(c, F.PRECISION_CHOSEN)
        F.plan_set_tuning(p, variant=v, split_m1=m1, chunk_frames=chunk)
        _compare(gpu_run(c, plan=p), o64, CHOSEN_TOL)
    with pytest.raises(F.FxcError):                   # the fused form needs M1 = 128
        F.plan_set_tuning(gpu_plan(c, F.PRECISION_CHOSEN), variant=1, split_m1=256,
                          chunk_frames=7)


@pytest.mark.parametrize("n,bits,sb,grid,prods,lsb,delay,nf,chunk", [
    (32768, (2, 2), 0, 0, "xa", (False, False), 5.3, 40, 0),
    (32768, (1, 8), 128, 0, "xa", (True, False), -3.6, 5, 0),        # fewer frames than groups
    (32768, (4, 2), -256, 1, "xa", (False, True), 17.2, 40, 7),      # several chunks
    (65536, (4, 1), 2048, 0, "ov", (True, True), 2.5, 24, 0),        # ACF on the overlap only
    (65536, (2, 4), -128, 1, "x", (False, False), -40.7, 24, 5),     # XCF only, odd shifts
])
def test_large_fused_form(n, bits, sb, grid, prods, lsb, delay, nf, chunk):
    """The fused four-step (decode in the first pass; the row DFTs, untangle
    and products in the second) against the FP64 oracle and against the
    four-step through HBM on the same plan: mixed code widths, band shifts
    (multiples of M1 = 128) on both output grids, LSB flips, integer shifts
    of either sign, fewer frames than frame groups and several chunks."""
    fs = 512e6
    band = (0, 0, n // 2, 0)
    if sb:
        band = (sb, 0, n // 2 - abs(sb), abs(sb)) if sb > 0 else (sb, abs(sb), n // 2, 0)
    products = {"xa": F.PRODUCT_ACF | F.PRODUCT_XCF, "x": F.PRODUCT_XCF,
                "ov": F.PRODUCT_ACF | F.PRODUCT_ACF_OVERLAP_ONLY | F.PRODUCT_XCF}[prods]
    c = make_case(n=n, fs=fs, nf=nf, bits=bits, lsb=lsb, band=band, grid=grid,
                  products=products, rho=0.4, delay_samples=delay, seed=n + nf + abs(sb),
                  delay_user_samples=delay, extra_delay_rate_sps=2e-9, lo1_hz=8.4e9,
                  lo2_hz=8.4e9 - sb * fs / n)
    o64 = oracle_run(c, R.MODE_F64)
    keys = ("acc11", "acc22", "acc12") if products & F.PRODUCT_ACF else ("acc12",)
    p = gpu_plan(c, F.PRECISION_CHOSEN)
    assert p.info.fft_form == 2 and p.info.variant == 1
    if chunk:
        F.plan_set_tuning(p, variant=1, split_m1=128, chunk_frames=chunk)
    fused = gpu_run(c, plan=p)
    _compare(fused, o64, CHOSEN_TOL, keys)
    q = gpu_plan(c, F.PRECISION_CHOSEN)
    F.plan_set_tuning(q, variant=0, split_m1=n // 2 // 128, chunk_frames=chunk or nf)
    hbm = gpu_run(c, plan=q)
    _compare(fused, hbm, CHOSEN_TOL, keys)
    for k in keys:
        assert zero_pattern_equal(fused[k], hbm[k]), k


@pytest.mark.parametrize("bits,sb,grid,prods,lsb,delay,nf,chunk,shuffle", [
    ((2, 2), 0, 0, "xa", (False, False), 5.3, 300, 0, False),
    ((1, 8), 128, 0, "xa", (True, False), -3.6, 5, 0, False),      # fewer frames than groups
    ((4, 2), -256, 1, "xa", (False, True), 17.2, 300, 37, True),   # several chunks, shuffle
    ((4, 1), 2048, 0, "ov", (True, True), 2.5, 200, 0, False),     # ACF on the overlap only
    ((2, 4), -128, 1, "x", (False, False), -40.7, 100, 23, False),  # XCF only, odd shifts
    ((3, 5), 0, 0, "xa", (True, False), 1.7, 120, 0, True),         # codes straddling words
])
def test_fused_four_step_fft16384(bits, sb, grid, prods, lsb, delay, nf, chunk, shuffle):
    """The fused four-step at N = 16384 (M1 = 128 x M2 = 64, the chosen
    form's default there; rows by two radix-8 passes) against the FP64
    oracle and against the Stockham kernel on the same plan (variant 0)."""
    n, fs = 16384, 512e6
    band = (0, 0, n // 2, 0)
    if sb:
        band = (sb, 0, n // 2 - abs(sb), abs(sb)) if sb > 0 else (sb, abs(sb), n // 2, 0)
    products = {"xa": F.PRODUCT_ACF | F.PRODUCT_XCF, "x": F.PRODUCT_XCF,
                "ov": F.PRODUCT_ACF | F.PRODUCT_ACF_OVERLAP_ONLY | F.PRODUCT_XCF}[prods]
    rng = np.random.default_rng(nf + abs(sb))
    c = make_case(n=n, fs=fs, nf=nf, bits=bits, lsb=lsb, band=band, grid=grid,
                  shuffles=(rng.permutation(32) if shuffle else None, None),
                  products=products, rho=0.4, delay_samples=delay, seed=n + nf + abs(sb),
                  delay_user_samples=delay, extra_delay_rate_sps=2e-9, lo1_hz=8.4e9,
                  lo2_hz=8.4e9 - sb * fs / n)
    o64 = oracle_run(c, R.MODE_F64)
    keys = ("acc11", "acc22", "acc12") if products & F.PRODUCT_ACF else ("acc12",)
    p = gpu_plan(c, F.PRECISION_CHOSEN)
    assert p.info.fft_form == 0 and p.info.variant == 1
    if chunk:
        F.plan_set_tuning(p, variant=1, split_m1=128, chunk_frames=chunk)
    fused = gpu_run(c, plan=p)
    _compare(fused, o64, CHOSEN_TOL, keys)
    q = gpu_plan(c, F.PRECISION_CHOSEN)
    F.plan_set_tuning(q, variant=0)
    stockham = gpu_run(c, plan=q)
    _compare(fused, stockham, CHOSEN_TOL, keys)
    for k in keys:
        assert zero_pattern_equal(fused[k], stockham[k]), k


def _raw_code_diff(got, ref, n, bit_out, shuffle, levels, nf):
    dp = out_decode_plan(bit_out, shuffle, levels)
    bpf_o = n * bit_out // 8
    diff = 0
    for f in range(nf):
        seg = slice(f * bpf_o, (f + 1) * bpf_o)
        _, vg = dp.decode(got[seg], n, False)
        _, vr = dp.decode(ref[seg], n, False)
        diff += int(np.count_nonzero(vg != vr))
    return diff


@pytest.mark.parametrize("bits,sb,grid,lsb_out,raw,nf,chunk,shuffle", [
    ((2, 8), 0, 0, False, True, 300, 0, True),
    ((8, 1), 128, 1, True, True, 5, 0, False),        # fewer frames than groups
    ((4, 2), -256, 0, False, True, 300, 37, True),    # several chunks
    ((1, 8), 2048, 1, True, True, 200, 0, False),
    ((5, 3), 0, 0, True, True, 120, 0, True),         # 5-bit output: codes straddle words
    ((2, 2), -128, 1, False, False, 100, 23, False),  # power only, several chunks
    ((2, 1), 0, 1, False, True, 1100, 0, False),      # past the default chunk (1024 frames)
])
def test_phased_fused_four_step_fft16384(bits, sb, grid, lsb_out, raw, nf, chunk, shuffle):
    """The phased sum on the fused four-step at N = 16384 (the chosen form's
    default there when the band shift is a multiple of 128): |cb|^2 within
    the chosen bar of the oracle's FP64 phased branch and of the Stockham
    phased kernel (variant 0), .raw codes equal except at rare quantiser
    decision boundaries."""
    n, fs = 16384, 256e6
    half = n // 2
    a1s = max(0, -sb)
    band = (sb, a1s, min(half, half - sb), a1s + sb)
    rng = np.random.default_rng(nf + abs(sb) + bits[0])
    prods = F.PRODUCT_PHASED | (F.PRODUCT_RAW_OUT if raw else 0)
    c = make_case(n=n, fs=fs, nf=nf, bits=bits, shuffles=(rng.permutation(32) if shuffle else None, None),
                  lsb=(False, True), band=band, grid=grid, rho=0.4, delay_samples=-7.3,
                  seed=n + nf + abs(sb), delay_user_samples=-7.3, extra_delay_rate_sps=2e-9,
                  lo1_hz=6.1e9, lo2_hz=6.1e9 - sb * fs / n, products=prods)
    w = (0.8, 1.25)
    bit_out = max(bits)

    def plan(variant):
        q = F.Plan(n, bits=bits, levels=c["levels"], shuffles=c["shuffles"], lsb=c["lsb"], band=band,
                   df_hz=fs / n, products=prods, out_grid=grid, w=w, bit_out=bit_out,
                   output_lsb=lsb_out, precision=F.PRECISION_CHOSEN)
        assert q.info.fft_form == 0 and q.info.variant == 1
        if variant == 1 and chunk:
            F.plan_set_tuning(q, variant=1, split_m1=128, chunk_frames=chunk)
        elif variant == 0:
            F.plan_set_tuning(q, variant=0)
        return q

    cfg = F.make_delay_cfg(c["table"], **c["delay"])
    got = plan(1).sector_run(c["raw1"], c["raw2"], 0, nf, cfg)
    stk = plan(0).sector_run(c["raw1"], c["raw2"], 0, nf, cfg)
    ref = R.sector_run(_phased_desc(c, w, bit_out, lsb_out), c["raw1"], c["raw2"], 0, nf,
                       R.make_delay_cfg(c["table"], **c["delay"]), mode=R.MODE_F64, threads=8)
    assert got["n_failed"] == 0 and ref["n_failed"] == 0
    e_ref, e_stk = rel_err(got["acc_ph"], ref["acc_ph"]), rel_err(got["acc_ph"], stk["acc_ph"])
    print(f"[phased fused] bits {bits} sb {sb}: acc_ph vs oracle {e_ref:.2e}, vs Stockham {e_stk:.2e}",
          flush=True)
    assert e_ref < CHOSEN_TOL and e_stk < CHOSEN_TOL
    for k in ("acc11", "acc22", "acc12"):
        assert np.all(got[k] == 0)
    if raw:
        assert got["raw_out"].size == nf * n * bit_out // 8
        for other in (ref, stk):
            d = _raw_code_diff(got["raw_out"], other["raw_out"], n, bit_out, c["shuffles"][0],
                               c["levels"][0], nf)
            assert d <= 1e-4 * n * nf + 2, d


def test_fused_four_step_parameter_spans_fft16384():
    """A 9,000-frame sector at N = 16384: the fused four-step computes the
    frame parameters for spans of whole chunks (8,192 frames here), so the
    sector takes two spans.  ACF/XCF and the phased sum (power and 2-bit
    .raw) against the Stockham kernels on the same plans."""
    n, fs, nf = 16384, 1024e6, 9000
    rng = np.random.default_rng(16384 + nf)
    raw1 = rng.integers(0, 256, nf * n // 4, dtype=np.uint8)
    raw2 = rng.integers(0, 256, nf * n // 4, dtype=np.uint8)
    cfg = F.make_delay_cfg(np.zeros(2 + int(nf * n / fs)), frame_dt=n / fs, fs=fs,
                           delay_user_samples=6.7, extra_delay_rate_sps=3e-9, lo1_hz=7e9, lo2_hz=7e9)
    lv = [LEVELS_2BIT, LEVELS_2BIT]
    res = {}
    for v in (1, 0):
        p = F.Plan(n, levels=lv, df_hz=fs / n, precision=F.PRECISION_CHOSEN)
        q = F.Plan(n, levels=lv, df_hz=fs / n, products=F.PRODUCT_PHASED | F.PRODUCT_RAW_OUT,
                   w=(0.9, 1.1), bit_out=2, precision=F.PRECISION_CHOSEN)
        assert p.info.variant == 1 and q.info.variant == 1
        if v == 0:
            F.plan_set_tuning(p, variant=0)
            F.plan_set_tuning(q, variant=0)
        res[v] = (p.sector_run(raw1, raw2, 0, nf, cfg), q.sector_run(raw1, raw2, 0, nf, cfg))
        p.close()
        q.close()
    _compare(res[1][0], res[0][0], CHOSEN_TOL)
    assert rel_err(res[1][1]["acc_ph"], res[0][1]["acc_ph"]) < CHOSEN_TOL
    # 2-bit codes with the identity output map: compare the fields directly
    sh = np.array([0, 2, 4, 6], dtype=np.uint8)
    a, b = res[1][1]["raw_out"], res[0][1]["raw_out"]
    assert a.size == b.size == nf * n // 4
    d = int(np.count_nonzero(((a[:, None] >> sh) & 3) != ((b[:, None] >> sh) & 3)))
    assert d <= 1e-4 * n * nf + 2, d


@pytest.mark.parametrize("n,bits", [(32768, (3, 5)), (65536, (5, 3))])
def test_large_fused_odd_widths_and_shuffle(n, bits):
    """The fused four-step's run-time code width (3 and 5 bits: codes that
    straddle words, the generic decode instance) with a generic bit shuffle
    on station 1, against the FP64 oracle and the four-step through HBM."""
    fs = 256e6
    rng = np.random.default_rng(n + 7)
    c = make_case(n=n, fs=fs, nf=20, bits=bits, shuffles=(rng.permutation(32), None),
                  lsb=(True, False), rho=0.4, delay_samples=-6.4, seed=n + 3,
                  delay_user_samples=-6.4, extra_delay_rate_sps=1e-9, lo1_hz=6e9, lo2_hz=6e9)
    o64 = oracle_run(c, R.MODE_F64)
    p = gpu_plan(c, F.PRECISION_CHOSEN)
    assert p.info.fft_form == 2 and p.info.variant == 1
    fused = gpu_run(c, plan=p)
    _compare(fused, o64, CHOSEN_TOL)
    q = gpu_plan(c, F.PRECISION_CHOSEN)
    F.plan_set_tuning(q, variant=0, split_m1=n // 2 // 128, chunk_frames=20)
    _compare(fused, gpu_run(c, plan=q), CHOSEN_TOL)


@pytest.mark.parametrize("n,prec", [(4096, F.PRECISION_PLAIN_FP32), (1024, F.PRECISION_CHOSEN),
                                    (65536, F.PRECISION_CHOSEN)])
def test_autotune_picks_a_candidate(n, prec):
    fs = 256e6
    nf = {4096: 20000, 1024: 50000, 65536: 300}[n]
    s = F.Stream()
    bpf = n * 2 // 8
    r1, r2 = F.DeviceBuffer(nf * bpf), F.DeviceBuffer(nf * bpf)
    F.synth_station_device(r1.ptr, nf * n, 2, None, False, 9, 0, 0.2, 3, s.ptr)
    F.synth_station_device(r2.ptr, nf * n, 2, None, False, 9, 1, 0.2, 3, s.ptr)
    s.sync()
    p = F.Plan(n, levels=[LEVELS_2BIT, LEVELS_2BIT], df_hz=fs / n, precision=prec)
    cfg = F.make_delay_cfg(np.zeros(4), frame_dt=n / fs, fs=fs, delay_user_samples=3.3,
                           lo1_hz=3e9)
    rep = F.plan_autotune(p, r1.ptr, r2.ptr, nf, cfg, s.ptr)
    assert rep["candidates"] >= 1 and 0 < rep["best_ms"] <= rep["default_ms"]
    if n == 4096:
        assert rep["candidates"] == 6
    if n == 1024:
        assert rep["candidates"] == 3 and rep["variant"] in (0, 1, 2)
    assert p.info.variant == rep["variant"] and p.info.grid_divisor == rep["grid_divisor"]
    if n == 65536:
        # the fused form (variant 1) at three chunk sizes, then the four-step
        # through HBM at two splits and three chunk sizes
        assert rep["candidates"] == 9 and rep["variant"] in (0, 1)
        assert rep["split_m1"] in (128, 256) and rep["chunk_frames"] > 0
        assert rep["variant"] == 0 or rep["split_m1"] == 128
