"""Pins the oracle's phased-array branch (the write_raw path of process_frame,
main.rs:2222-2365: per-station rotation utils.rs:69-124, weighted sum and
band alignment main.rs:2264-2302, C2R and quantise_frame utils.rs:378-410)
in both modes -- the reference-faithful f32 one and the FP64 one the chosen
form is held to -- by cases whose answer the mathematics fixes, and by an
independent numpy FP64 re-run of the whole composition.

The inputs are codes drawn in numpy, packed by the decode contract
(_synth.pack_codes), so the exact station samples are levels[codes] without
the oracle's decoder."""
import numpy as np
import pytest

import fxref as R
from _synth import IDENT, LEVELS_2BIT, pack_codes, uniform_levels

FS = 64e6


def _cfg(n, table=None, **kw):
    base = dict(frame_dt=n / FS, fs=FS, time_offset_s=0.0, coarse_delay_s=0.0,
                delay_user_samples=0.0, extra_delay_rate_sps=0.0, clock1_delay_s=0.0,
                clock1_rate_sps=0.0, clock2_delay_s=0.0, clock2_rate_sps=0.0, d_seek=0.0,
                net_d1_base=0.0, total_rate1_base=0.0, total_accel1_base=0.0, net_d2_base=0.0,
                total_rate2_base=0.0, total_accel2_base=0.0, lo1_hz=0.0, lo2_hz=0.0)
    base.update(kw)
    return R.make_delay_cfg(np.zeros(8) if table is None else table, **base)


def _desc(n, bits=(2, 2), levels=(LEVELS_2BIT, LEVELS_2BIT), sh=(IDENT, IDENT), lsb=(0, 0),
          band=None, grid=0, w=(1.0, 1.0), bit_out=None, output_lsb=0):
    if band is None:
        band = (0, n // 2, 0)
    return dict(fft_len=n, bits1=bits[0], bits2=bits[1], levels1=levels[0], levels2=levels[1],
                shuffle1=sh[0], shuffle2=sh[1], lsb1=int(lsb[0]), lsb2=int(lsb[1]),
                a1s=band[0], a1e=band[1], a2s=band[2], out_grid=grid, df_hz=FS / n,
                need_acf=0, need_xcf=0, need_phased=1, write_raw=1, w1=w[0], w2=w[1],
                bit_out=bit_out or max(bits), output_lsb=int(output_lsb), cpu_threads=4)


def _codes(nf, n, bits, seed):
    return np.random.default_rng(seed).integers(0, 1 << bits, size=(nf, n))


def _acf64(x):
    """sum over frames of |rfft(frame)|^2 in FP64."""
    return np.sum(np.abs(np.fft.rfft(x, axis=1)) ** 2, axis=0)


MODES = [R.MODE_F32, R.MODE_F64]
TOL = {R.MODE_F32: 2e-6, R.MODE_F64: 1e-12}   # f32 FFT rounding / FP64 rounding


def _run(desc, raw1, raw2, nf, cfg, mode=R.MODE_F32):
    r = R.sector_run(desc, raw1, raw2, 0, nf, cfg, mode=mode, threads=4)
    assert r["n_failed"] == 0
    return r


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("w1", [1.0, 0.7, 2.5])
def test_station2_weight_zero_gives_weighted_acf(w1, mode):
    """w2 = 0, zero delay, no fringe: cb = w1 s1, so acc_ph = w1^2 ACF11."""
    n, nf = 256, 96
    c1, c2 = _codes(nf, n, 2, 1), _codes(nf, n, 2, 2)
    raw1, raw2 = pack_codes(c1, 2, n), pack_codes(c2, 2, n)
    r = _run(_desc(n, w=(w1, 0.0)), raw1, raw2, nf, _cfg(n), mode)
    ref = (w1 * w1) * _acf64(LEVELS_2BIT[c1])
    assert np.max(np.abs(r["acc_ph"] - ref) / ref) < TOL[mode]


@pytest.mark.parametrize("mode", MODES)
def test_identical_stations_give_four_times_acf(mode):
    """Both stations the same bytes, zero delay, w1 = w2 = 1: cb = 2 s on the
    overlap [0, N/2) and s at the Nyquist bin (outside the overlap, so only
    station 1 contributes there): acc_ph = 4 ACF on the overlap, ACF at N/2.
    Doubling is exact in f32, so the relation holds bit for bit against the
    station-1-only run."""
    n, nf = 512, 64
    c = _codes(nf, n, 2, 3)
    raw = pack_codes(c, 2, n)
    both = _run(_desc(n, w=(1.0, 1.0)), raw, raw, nf, _cfg(n), mode)
    one = _run(_desc(n, w=(1.0, 0.0)), raw, raw, nf, _cfg(n), mode)
    assert np.array_equal(both["acc_ph"][: n // 2], 4.0 * one["acc_ph"][: n // 2])
    assert both["acc_ph"][n // 2] == one["acc_ph"][n // 2]
    ref = _acf64(LEVELS_2BIT[c])
    assert np.max(np.abs(one["acc_ph"] - ref) / ref) < TOL[mode]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("bits,sh_seed,output_lsb", [(2, None, False), (2, 7, False), (4, 9, False),
                                                     (2, None, True), (8, 11, True), (1, 5, False)])
def test_raw_round_trip_is_exact(bits, sh_seed, output_lsb, mode):
    """w1 = 1, w2 = 0, zero delay and fringe, bit_out = station-1 bits: the
    time series is station 1's samples up to FFT round-trip rounding, far
    below half a level step, so the quantiser returns the input codes and
    the .raw bytes equal station 1's bytes, shuffle included.  With
    output_lsb the odd samples are negated first: for levels symmetric about
    zero that maps code k to 2^bits - 1 - k on odd samples only."""
    n, nf = 256, 40
    lv = LEVELS_2BIT if bits == 2 else uniform_levels(bits)
    sh = IDENT if sh_seed is None else np.random.default_rng(sh_seed).permutation(32).astype(np.int32)
    c1, c2 = _codes(nf, n, bits, 21), _codes(nf, n, 2, 22)
    raw1, raw2 = pack_codes(c1, bits, n, sh), pack_codes(c2, 2, n)
    d = _desc(n, bits=(bits, 2), levels=(lv, LEVELS_2BIT), sh=(sh, IDENT), w=(1.0, 0.0),
              bit_out=bits, output_lsb=output_lsb)
    r = _run(d, raw1, raw2, nf, _cfg(n), mode)
    if not output_lsb:
        assert np.array_equal(r["raw_out"], raw1)
    else:
        exp = c1.copy()
        exp[:, 1::2] = (1 << bits) - 1 - exp[:, 1::2]
        assert np.array_equal(r["raw_out"], pack_codes(exp, bits, n, sh))
        assert not np.array_equal(r["raw_out"], raw1)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("sb", [8, -8, 40])
def test_band_shift_places_station2_on_overlap_bins(sb, mode):
    """Station 2 alone (w1 = 0) on the Ant1 grid lands on grid bins
    [a1s, a1e) only, each carrying station 2's bin a2s + (k - a1s); on the
    Ant2 grid the roles swap: station 1 alone (the aligned side, weight w1)
    lands on [a2s, a2s + ov), and station 2 alone covers every bin."""
    n, nf = 256, 48
    half = n // 2
    a1s = max(0, -sb)
    a1e = min(half, half - sb)
    a2s = a1s + sb
    ov = a1e - a1s
    c1, c2 = _codes(nf, n, 2, 31), _codes(nf, n, 2, 32)
    raw1, raw2 = pack_codes(c1, 2, n), pack_codes(c2, 2, n)
    acf1, acf2 = _acf64(LEVELS_2BIT[c1]), _acf64(LEVELS_2BIT[c2])
    band = (a1s, a1e, a2s)
    cfg = _cfg(n)

    r = _run(_desc(n, band=band, grid=0, w=(0.0, 1.0)), raw1, raw2, nf, cfg, mode)
    m = np.zeros(half + 1, bool)
    m[a1s:a1e] = True
    assert np.all(r["acc_ph"][~m] == 0)
    ref = acf2[a2s:a2s + ov]
    assert np.max(np.abs(r["acc_ph"][m] - ref) / ref) < TOL[mode]

    r = _run(_desc(n, band=band, grid=1, w=(1.0, 0.0)), raw1, raw2, nf, cfg, mode)
    m = np.zeros(half + 1, bool)
    m[a2s:a2s + ov] = True
    assert np.all(r["acc_ph"][~m] == 0)
    ref = acf1[a1s:a1e]
    assert np.max(np.abs(r["acc_ph"][m] - ref) / ref) < TOL[mode]

    r = _run(_desc(n, band=band, grid=1, w=(0.0, 1.0)), raw1, raw2, nf, cfg, mode)
    assert np.max(np.abs(r["acc_ph"] - acf2) / acf2) < TOL[mode]


def _numpy_phased(c1, c2, lv, n, nf, band, grid, w, cfg, lsb, output_lsb, bit_out, lv_out,
                  fringe64=False):
    """The phased composition in FP64 from its definition: samples, shift
    with zero fill, rfft, per-station fractional-delay rotation exp(-2 pi i
    df frac k) with the Nyquist imaginary part cleared, fringe phasor, band
    alignment onto the output grid, weighted sum, |cb|^2; then DC/Nyquist
    imaginary parts cleared, irfft (1/N), odd samples negated for
    output_lsb, and the first nearest level of station 1's table."""
    half = n // 2 + 1
    a1s, a1e, a2s = band
    ov = a1e - a1s
    df = FS / n
    k = np.arange(half)
    acc = np.zeros(half)
    codes = np.zeros((nf, n), dtype=np.int64)
    xs = []
    for st, c in enumerate((c1, c2)):
        x = lv[st][c].astype(np.float64)
        if lsb[st]:
            x[:, 1::2] *= -1.0
        xs.append(x)
    lvo = np.asarray(lv_out, dtype=np.float64)
    for f in range(nf):
        fd = R.frame_delay(f, cfg)
        fr64 = R.frame_delay_f64(f, cfg)[1:] if fringe64 else None
        spec = []
        for st, (x, s, frac, fr) in enumerate(((xs[0][f], fd.int1, fd.frac1, fd.fr_lo1),
                                               (xs[1][f], fd.int2, fd.frac2, fd.fr_lo2))):
            y = np.zeros(n)
            if abs(s) < n:
                if s >= 0:
                    y[s:] = x[: n - s]
                else:
                    y[: n + s] = x[-s:]
            X = np.fft.rfft(y) * np.exp(-2j * np.pi * df * frac * k)
            X[n // 2] = X[n // 2].real
            spec.append(X * (fr64[st] if fringe64 else complex(fr[0], fr[1])))
        cb = np.zeros(half, complex)
        if grid == 0:
            cb += w[0] * spec[0]
            cb[a1s:a1e] += w[1] * spec[1][a2s:a2s + ov]
        else:
            cb += w[1] * spec[1]
            cb[a2s:a2s + ov] += w[0] * spec[0][a1s:a1e]
        acc += np.abs(cb) ** 2
        cb[0] = cb[0].real
        cb[n // 2] = cb[n // 2].real
        t = np.fft.irfft(cb, n)
        if output_lsb:
            t[1::2] *= -1.0
        codes[f] = np.argmin(np.abs(t[:, None] - lvo[None, :]), axis=1)
    return acc, codes, t


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("n,grid,sb,lsb,output_lsb", [
    (256, 0, 0, (0, 0), False),
    (512, 1, 16, (0, 1), True),
    (256, 0, -24, (1, 0), False),
    (1024, 1, 0, (1, 1), True),
])
def test_phased_composition_vs_numpy_fp64(n, grid, sb, lsb, output_lsb, mode):
    """The oracle's phased branch against the FP64 definition above with a
    delay of several samples (integer shift + fraction), a delay rate and
    fringe phasors of 5 GHz LOs.  f32 mode (fringe phasors the reference's
    f32 values): |cb|^2 within 5e-5 (f32 FFT and the f32 rotation
    recurrence), .raw codes equal except where the FP64 series sits within
    a hair of a decision boundary (neighbours only).  FP64 mode (FP64
    fringe phasors): |cb|^2 within 1e-10, codes equal except at most a few
    decided by the f32 rounding of the quantiser's input."""
    nf = 48
    half = n // 2
    a1s = max(0, -sb)
    band = (a1s, min(half, half - sb), a1s + sb)
    rng = np.random.default_rng(n + sb)
    # correlated Gaussian pair quantised to 2 bits, so the combined series is not trivial
    common = rng.standard_normal((nf, n))
    x1 = 0.6 * common + 0.8 * rng.standard_normal((nf, n))
    x2 = 0.6 * np.roll(common, 3, axis=1) + 0.8 * rng.standard_normal((nf, n))
    q = lambda x: np.digitize(x, [-0.98, 0.0, 0.98])
    c1, c2 = q(x1), q(x2)
    sh1 = rng.permutation(32).astype(np.int32)
    raw1, raw2 = pack_codes(c1, 2, n, sh1), pack_codes(c2, 2, n)
    w = (0.8, 1.3)
    kw = dict(delay_user_samples=3.3, extra_delay_rate_sps=2e-7, lo1_hz=5e9, lo2_hz=5.1e9)
    cfg = _cfg(n, **kw)
    d = _desc(n, sh=(sh1, IDENT), lsb=lsb, band=band, grid=grid, w=w, bit_out=2,
              output_lsb=output_lsb)
    r = _run(d, raw1, raw2, nf, cfg, mode)
    f64 = mode == R.MODE_F64
    acc, codes, _ = _numpy_phased(c1, c2, (LEVELS_2BIT, LEVELS_2BIT), n, nf, band, grid, w, cfg,
                                  lsb, output_lsb, 2, LEVELS_2BIT, fringe64=f64)
    assert np.max(np.abs(r["acc_ph"] - acc) / acc) < (1e-10 if f64 else 5e-5)
    got = np.zeros((nf, n), dtype=np.int64)
    dp = R.DecodePlan(2, sh1, LEVELS_2BIT)
    bpf = n * 2 // 8
    for f in range(nf):
        ok, v = dp.decode(r["raw_out"][f * bpf:(f + 1) * bpf], n, False)
        assert ok
        got[f] = np.searchsorted(LEVELS_2BIT, v)
    diff = got != codes
    assert np.all(np.abs(got - codes)[diff] == 1)
    assert np.count_nonzero(diff) <= (2 if f64 else 1e-3 * nf * n), np.count_nonzero(diff)
    # a planted mistake in the composition is caught: the weights swapped
    acc_bad, _, _ = _numpy_phased(c1, c2, (LEVELS_2BIT, LEVELS_2BIT), n, nf, band, grid,
                                  (w[1], w[0]), cfg, lsb, output_lsb, 2, LEVELS_2BIT,
                                  fringe64=f64)
    assert np.max(np.abs(r["acc_ph"] - acc_bad) / acc_bad) > 1e-2
