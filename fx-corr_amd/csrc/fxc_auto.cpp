// Auto-spectrum finalisation (acf.rs:3-23) and the numbers behind the
// phased-array plots of yi-phasedarray (main.rs:2727-2768): normalised
// power spectra with DC zeroed, their amplitudes on the observing-frequency
// axis, and the phased autocorrelation, i.e. the inverse complex FFT (1/N)
// of the spectrum mirrored into N bins, rotated so lag 0 sits at index N/2.
// Drawing the PNGs is left to the caller.
#include <cmath>
#include <complex>
#include <vector>

#include "fxcorr.h"

extern "C" fxc_status fxc_internal_fail(fxc_status s, const char *msg);

#pragma clang fp contract(off)

namespace {

// In-place inverse DFT, exp(+2 pi i j k / n), unnormalised, double precision:
// iterative radix-2 for powers of two, the direct sum otherwise.
void inverse_dft(std::vector<std::complex<double>> &x) {
    const size_t n = x.size();
    if (n < 2) return;
    if ((n & (n - 1)) == 0) {
        for (size_t i = 1, j = 0; i < n; ++i) {   // bit-reversal permutation
            size_t bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j |= bit;
            if (i < j) std::swap(x[i], x[j]);
        }
        for (size_t len = 2; len <= n; len <<= 1) {
            const size_t h = len / 2;
            for (size_t k = 0; k < h; ++k) {
                const double a = 2.0 * M_PI * (double)k / (double)len;
                const std::complex<double> w(std::cos(a), std::sin(a));
                for (size_t s = 0; s < n; s += len) {
                    const std::complex<double> u = x[s + k], v = x[s + k + h] * w;
                    x[s + k] = u + v;
                    x[s + k + h] = u - v;
                }
            }
        }
        return;
    }
    std::vector<std::complex<double>> y(n);
    for (size_t k = 0; k < n; ++k) {
        std::complex<double> acc(0.0, 0.0);
        for (size_t j = 0; j < n; ++j) {
            const double a = 2.0 * M_PI * (double)((j * k) % n) / (double)n;
            acc += x[j] * std::complex<double>(std::cos(a), std::sin(a));
        }
        y[k] = acc;
    }
    x.swap(y);
}

}  // namespace

extern "C" fxc_status fxc_finalize_auto_spectrum(const double *acc, int64_t len, double power_norm, double *out) {
    if (!acc || !out || len < 0) return fxc_internal_fail(FXC_ERR_INVALID_ARG, "bad finalize_auto_spectrum arguments");
    for (int64_t i = 0; i < len; ++i) out[i] = power_norm > 0.0 ? acc[i] / power_norm : acc[i];
    if (len > 0) out[0] = 0.0;
    return FXC_OK;
}

extern "C" fxc_status fxc_phased_plot_series(const double *acc_ph, const double *acc11, const double *acc22,
                                             int64_t len, int32_t fft_len, int64_t emitted_frames, double obs_mhz,
                                             double df_hz, fxc_phased_plot *out) {
    if (!acc_ph || !acc11 || !acc22 || !out || fft_len < 1 || len != fft_len / 2 + 1 || emitted_frames < 0)
        return fxc_internal_fail(FXC_ERR_INVALID_ARG, "bad phased plot arguments");
    const int64_t bins = (fft_len % 2 == 0) ? fft_len / 2 : len;   // Nyquist bin left out for even N
    out->spec_bins = bins;
    if (!out->freqs_mhz || !out->phased_power || !out->amp_ph || !out->amp11 || !out->amp22 || !out->acf_shifted)
        return fxc_internal_fail(FXC_ERR_INVALID_ARG, "phased plot output arrays missing");
    const double norm = std::fmax((double)emitted_frames * (double)fft_len, 1.0);
    std::vector<double> ph(len), a11(len), a22(len);
    fxc_finalize_auto_spectrum(acc_ph, len, norm, ph.data());
    fxc_finalize_auto_spectrum(acc11, len, norm, a11.data());
    fxc_finalize_auto_spectrum(acc22, len, norm, a22.data());
    const double df = df_hz / 1e6;
    for (int64_t i = 0; i < bins; ++i) {
        out->freqs_mhz[i] = obs_mhz + (double)i * df;
        out->phased_power[i] = ph[i];
        out->amp_ph[i] = std::sqrt(ph[i]);
        out->amp11[i] = std::sqrt(a11[i]);
        out->amp22[i] = std::sqrt(a22[i]);
    }
    // the spectrum as f32, mirrored into bins N-i for 0 < i < N/2 (for odd N
    // bin N - N/2 stays zero, as in the reference)
    std::vector<std::complex<double>> full(fft_len, 0.0);
    for (int64_t i = 0; i < bins; ++i) {
        const double v = (double)(float)ph[i];
        full[i] = v;
        if (i > 0 && i < fft_len / 2) full[fft_len - i] = v;
    }
    inverse_dft(full);
    const double inv_n = 1.0 / (double)fft_len;
    for (int32_t i = 0; i < fft_len; ++i)
        out->acf_shifted[i] = (double)(float)(full[(i + fft_len / 2) % fft_len].real() * inv_n);
    return FXC_OK;
}
