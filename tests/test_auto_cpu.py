"""finalize_auto_spectrum (acf.rs:3-23) and the phased-array plot series
(main.rs:2727-2768) on the CPU.  The oracle is pinned by closed forms (a
single spectral line gives a cosine autocorrelation; lag 0 is the mean of
the mirrored spectrum; DC zeroing and the power normalisation) and by
numpy's inverse FFT; the library is held to the oracle bit for bit on the
spectra and within one f32 rounding on the autocorrelation (both evaluate
the inverse DFT in FP64 and round the result to f32, as the reference's
f32 transform returns f32)."""
import math

import numpy as np
import pytest

import fxcorr as F
import fxref as R


def test_oracle_finalize_auto_spectrum():
    acc = np.array([5.0, 2.0, 4.0, 8.0])
    assert list(R.finalize_auto_spectrum(acc, 2.0)) == [0.0, 1.0, 2.0, 4.0]
    assert list(R.finalize_auto_spectrum(acc, 0.0)) == [0.0, 2.0, 4.0, 8.0]
    assert list(R.finalize_auto_spectrum(acc, -3.0)) == [0.0, 2.0, 4.0, 8.0]
    assert R.finalize_auto_spectrum(np.zeros(0), 1.0).size == 0


@pytest.mark.parametrize("n,j0", [(1024, 37), (4096, 1), (1000, 250), (64, 31)])
def test_oracle_single_line_gives_cosine_acf(n, j0):
    h = n // 2 + 1
    emitted = 3
    a = 2.5
    acc = np.zeros(h)
    acc[j0] = a * emitted * n                     # normalises to a
    acc[0] = 123.0                                # DC is zeroed
    r = R.phased_plot_series(acc, acc, acc, n, emitted, 6600.0, 15625.0)
    k = np.arange(n) - n // 2                     # lag of each acf_shifted entry
    want = 2.0 * a / n * np.cos(2 * np.pi * j0 * k / n)
    np.testing.assert_allclose(r["acf_shifted"], want, rtol=0, atol=2e-7 * 2 * a / n + 1e-12)
    assert r["phased_power"][j0] == a and r["phased_power"][0] == 0.0
    assert r["amp_ph"][j0] == math.sqrt(a)
    assert r["freqs_mhz"][10] == 6600.0 + 10 * (15625.0 / 1e6)
    assert r["phased_power"].size == n // 2      # even N: Nyquist bin not plotted


def test_oracle_lag_zero_is_spectrum_mean():
    n = 2048
    rng = np.random.default_rng(3)
    acc = rng.uniform(0.5, 2.0, n // 2 + 1)
    r = R.phased_plot_series(acc, acc, acc, n, 0, 0.0, 1.0)   # emitted 0 -> norm 1
    ph = r["phased_power"].astype(np.float32).astype(np.float64)
    lag0 = (ph[0] + 2.0 * ph[1:n // 2].sum()) / n
    assert r["acf_shifted"][n // 2] == pytest.approx(lag0, rel=1e-6)
    # symmetric spectrum -> even autocorrelation
    acf = r["acf_shifted"]
    np.testing.assert_allclose(acf[n // 2 + 1:], acf[n // 2 - 1:0:-1], rtol=0, atol=1e-6 * abs(acf).max())


@pytest.mark.parametrize("n", [1024, 4096, 65536, 1000, 7, 2])
def test_library_matches_oracle(n):
    rng = np.random.default_rng(n)
    h = n // 2 + 1
    accs = [rng.uniform(0.0, 1e9, h) for _ in range(3)]
    got = F.phased_plot_series(*accs, n, 5000, 6700.0, 512e6 / n)
    ora = R.phased_plot_series(*accs, n, 5000, 6700.0, 512e6 / n)
    for k in ("freqs_mhz", "phased_power", "amp_ph", "amp11", "amp22"):
        assert np.array_equal(got[k], ora[k]), k
    np.testing.assert_allclose(got["acf_shifted"], ora["acf_shifted"], rtol=0,
                               atol=1.2e-7 * np.abs(ora["acf_shifted"]).max())
    for a in accs:
        assert np.array_equal(F.finalize_auto_spectrum(a, 3.5e6), R.finalize_auto_spectrum(a, 3.5e6))


def test_odd_length_mirror_rule():
    """Odd N: bins 1 .. N/2-1 are mirrored, bin N/2 is not (main.rs:2755-2759)."""
    n, emitted = 7, 1
    acc = np.array([0.0, 7.0, 14.0, 21.0])        # -> 1, 2, 3 after / (1 * 7)
    r = F.phased_plot_series(acc, acc, acc, n, emitted, 0.0, 1.0)
    assert r["phased_power"].size == 4
    full = np.array([0.0, 1.0, 2.0, 3.0, 0.0, 2.0, 1.0])
    want = np.roll(np.fft.ifft(full).real, -(n // 2))
    np.testing.assert_allclose(r["acf_shifted"], want, rtol=0, atol=1e-7)


def test_plot_series_rejects_bad_lengths():
    with pytest.raises(F.FxcError):
        F.phased_plot_series(np.zeros(10), np.zeros(10), np.zeros(10), 1024, 1, 0.0, 1.0)
