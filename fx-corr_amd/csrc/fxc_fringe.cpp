// Whole-scan cross spectrum (the reference's --fringe / XCF pipeline):
// per-chunk XCF sums scaled by 1/N^2 into one FP64 spectrum
// (main.rs:1692-1909), then finalize_cross_spectrum (xcf.rs:13-104):
// DC / N/2 zeroed, magnitude and phase per bin, phase unwrapped from bin 1
// on, and a least-squares line through phase against sky frequency over
// bins 1 .. N/2-1 whose slope gives the delay.  The plots of that function
// are out of scope; the numbers they are drawn from are returned.
#include <cmath>
#include <cstring>
#include <vector>

#include "fxcorr.h"

extern "C" fxc_status fxc_internal_fail(fxc_status s, const char *msg);

// utils.rs:425 safe_arg: arg of 0 is 0.
static double safe_arg(double re, double im) { return (re == 0.0 && im == 0.0) ? 0.0 : std::atan2(im, re); }

extern "C" fxc_status fxc_finalize_cross_spectrum(double *spec, int64_t len, int32_t fft_len,
                                                  double sampling_rate_hz, double sky_low_mhz,
                                                  int32_t is_lsb, fxc_cross_spectrum_result *res,
                                                  double *mag_out, double *phase_out,
                                                  double *unwrapped_out, double *fit_out) {
#pragma clang fp contract(off)   // Rust evaluates every product and sum separately
    if (!spec || !res || len < 1 || fft_len < 1 || !(sampling_rate_hz > 0.0))
        return fxc_internal_fail(FXC_ERR_INVALID_ARG, "bad finalize_cross_spectrum arguments");
    const size_t n = (size_t)len;
    // xcf.rs:26-30: zero DC (index 0) and index fft_len/2 when it is inside
    spec[0] = 0.0;
    spec[1] = 0.0;
    if ((size_t)(fft_len / 2) < n) {
        spec[2 * (fft_len / 2)] = 0.0;
        spec[2 * (fft_len / 2) + 1] = 0.0;
    }
    // xcf.rs:32-42: magnitude / phase; the DC magnitude is replaced by the
    // mean of the others halved
    std::vector<double> mag(n), ph(n);
    for (size_t i = 0; i < n; ++i) {
        mag[i] = std::hypot(spec[2 * i], spec[2 * i + 1]);
        ph[i] = safe_arg(spec[2 * i], spec[2 * i + 1]);
    }
    if (n > 1) {
        double rest = 0.0;
        for (size_t i = 1; i < n; ++i) rest += mag[i];
        mag[0] = rest / (double)(n - 1) / 2.0;
    }
    // xcf.rs:44-53 with utils.rs:412-424: bin 0 kept, bins 1.. unwrapped
    // among themselves (offset += / -= 2 pi where the step exceeds pi)
    std::vector<double> unw(ph);
    if (n > 2) {
        double offset = 0.0;
        for (size_t i = 2; i < n; ++i) {
            const double diff = ph[i] - ph[i - 1];
            if (diff > M_PI) offset -= 2.0 * M_PI;
            else if (diff < -M_PI) offset += 2.0 * M_PI;
            unw[i] = ph[i] + offset;
        }
    }
    // xcf.rs:55-63: sky frequency of each bin (Hz)
    const double bw_mhz = sampling_rate_hz / 2.0 / 1e6;
    const double df_hz = sampling_rate_hz / (double)fft_len;
    std::vector<double> fhz(n);
    for (size_t i = 0; i < n; ++i) {
        const double fmhz = is_lsb ? (sky_low_mhz + bw_mhz) - ((double)i * df_hz) / 1e6
                                   : sky_low_mhz + ((double)i * df_hz) / 1e6;
        fhz[i] = fmhz * 1.0e6;
    }
    // xcf.rs:65-101: least squares over [1, end), end = len - 1 for even N
    res->has_delay = 0;
    res->delay_seconds = 0.0;
    res->slope_rad_per_hz = 0.0;
    res->intercept_rad = 0.0;
    bool fitted = false;
    if (n > 1) {
        const size_t start = 1;
        size_t end = n;
        if (fft_len % 2 == 0 && end > 1) end -= 1;
        if (end > start) {
            const double cnt = (double)(end - start);
            double sx = 0.0, sy = 0.0, sxy = 0.0, sx2 = 0.0;
            for (size_t i = start; i < end; ++i) sx += fhz[i];
            for (size_t i = start; i < end; ++i) sy += unw[i];
            for (size_t i = start; i < end; ++i) sxy += fhz[i] * unw[i];
            for (size_t i = start; i < end; ++i) sx2 += fhz[i] * fhz[i];
            const double den = cnt * sx2 - sx * sx;
            if (std::fabs(den) > 1e-9) {
                const double slope = (cnt * sxy - sx * sy) / den;
                const double icpt = (sy - slope * sx) / cnt;
                res->has_delay = 1;
                res->delay_seconds = -slope / (2.0 * M_PI);
                res->slope_rad_per_hz = slope;
                res->intercept_rad = icpt;
                fitted = true;
            }
        }
    }
    for (size_t i = 0; i < n; ++i) {
        if (mag_out) mag_out[i] = mag[i];
        if (phase_out) phase_out[i] = ph[i];
        if (unwrapped_out) unwrapped_out[i] = unw[i];
        if (fit_out) fit_out[i] = fitted ? res->slope_rad_per_hz * fhz[i] + res->intercept_rad : 0.0;
    }
    return FXC_OK;
}
