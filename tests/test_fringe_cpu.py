"""finalize_cross_spectrum (xcf.rs:13-104) on the CPU: the oracle pinned by
closed forms (a linear phase of known delay, on USB and LSB frequency axes,
including slopes that need unwrapping; the DC / N/2 zeroing and the DC
magnitude rule; degenerate lengths) and by numpy's unwrap and polyfit; then
the C-ABI function against the oracle on random spectra."""
import math

import numpy as np
import pytest

import fxcorr as F
import fxref as R


def _linear_phase(n_bins, fft_len, fs, sky_mhz, tau, phi0, is_lsb, amp=None):
    df = fs / fft_len
    i = np.arange(n_bins)
    f = (sky_mhz + fs / 2e6 - i * df / 1e6) if is_lsb else (sky_mhz + i * df / 1e6)
    a = np.ones(n_bins) if amp is None else amp
    return a * np.exp(1j * (phi0 - 2.0 * np.pi * tau * f * 1e6)), f * 1e6


@pytest.mark.parametrize("is_lsb", [False, True])
@pytest.mark.parametrize("tau_samples", [0.0, 3.7, -41.2, 380.0])
def test_oracle_recovers_linear_phase_delay(is_lsb, tau_samples):
    fft_len, fs, sky = 1024, 64e6, 6700.0
    tau = tau_samples / fs                       # per-bin phase step 2 pi tau df < pi
    amp = 1.0 + 0.3 * np.cos(np.arange(513) / 7.0)
    spec, _ = _linear_phase(513, fft_len, fs, sky, tau, 0.4, is_lsb, amp)
    out, r = R.finalize_cross_spectrum(spec, fft_len, fs, sky, is_lsb)
    assert out[0] == 0 and out[512] == 0                          # xcf.rs:26-30
    assert r["mag"][0] == pytest.approx(np.mean(r["mag"][1:]) / 2.0, rel=1e-15)
    assert r["mag"][512] == 0.0
    assert r["delay"] == pytest.approx(tau, rel=1e-7, abs=1e-18)
    # numpy: the tail unwrap and a least-squares line over bins 1..511
    np.testing.assert_allclose(r["unwrapped"][1:], np.unwrap(r["phase"][1:]), rtol=0, atol=1e-12)
    slope = np.polyfit(r["fhz"][1:512], r["unwrapped"][1:512], 1)[0]
    assert -slope / (2 * math.pi) == pytest.approx(r["delay"], rel=1e-9, abs=1e-18)


def test_oracle_degenerate_lengths():
    _, r = R.finalize_cross_spectrum(np.array([1 + 1j]), 4, 1e6, 100.0)
    assert r["delay"] is None and r["mag"][0] == 0.0
    _, r = R.finalize_cross_spectrum(np.array([1 + 1j, 2 - 1j]), 2, 1e6, 100.0)
    assert r["delay"] is None                    # even N: fit range [1, 1) is empty
    out, r = R.finalize_cross_spectrum(np.array([1 + 1j, 2 - 1j]), 3, 1e6, 100.0)
    assert r["delay"] is None                    # odd N, one point: n sx2 - sx^2 = 0
    assert out[1] == 0                           # fft_len/2 = 1 is zeroed
    out, r = R.finalize_cross_spectrum(np.array([1 + 1j, 2 - 1j, 1j]), 3, 1e6, 100.0)
    assert r["delay"] is not None                # odd N fits bins 1..len-1: two points


@pytest.mark.parametrize("seed,fft_len,is_lsb", [(1, 1024, False), (2, 1024, True),
                                                 (3, 65536, False), (4, 1023, True)])
def test_finalize_matches_oracle(seed, fft_len, is_lsb):
    rng = np.random.default_rng(seed)
    n = fft_len // 2 + 1
    tau = rng.uniform(-30, 30) / 512e6
    spec, _ = _linear_phase(n, fft_len, 512e6, 8400.0, tau, 1.1, is_lsb)
    spec = spec * (1 + 0.2 * rng.standard_normal(n)) + 0.05 * (rng.standard_normal(n)
                                                               + 1j * rng.standard_normal(n))
    spec[17] = 0.0                               # safe_arg of zero
    o_spec, o = R.finalize_cross_spectrum(spec, fft_len, 512e6, 8400.0, is_lsb)
    g_spec, g = F.finalize_cross_spectrum(spec, fft_len, 512e6, 8400.0, is_lsb)
    assert np.array_equal(g_spec, o_spec)
    # glibc hypot (the product, as Rust's norm) vs Python's own: within 1 ulp
    np.testing.assert_allclose(g["mag"], o["mag"], rtol=5e-16, atol=0)
    assert np.array_equal(g["phase"], o["phase"])
    assert np.array_equal(g["unwrapped"], o["unwrapped"])
    assert g["delay"] == pytest.approx(o["delay"], rel=1e-12)
    assert g["slope"] == pytest.approx(o["slope"], rel=1e-12)
    np.testing.assert_allclose(g["fit"], o["slope"] * o["fhz"] + o["intercept"], rtol=1e-12,
                               atol=1e-9)


def test_finalize_rejects_bad_arguments():
    with pytest.raises(F.FxcError):
        F.finalize_cross_spectrum(np.zeros(0), 1024, 64e6, 100.0)
    with pytest.raises(F.FxcError):
        F.finalize_cross_spectrum(np.ones(4), 8, 0.0, 100.0)
