/*
 * fxref.c — CPU oracle (test infrastructure only; see fxref.h).
 *
 * Built with -ffp-contract=off -fno-fast-math so every f32/f64 operation is
 * rounded individually, in the order the reference's Rust writes it.
 */
#include "fxref.h"

#include <ctype.h>
#include <errno.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define FXR_PI 3.14159265358979323846

/* ------------------------------------------------------------------------ */
/* small helpers                                                            */
/* ------------------------------------------------------------------------ */

/* Rust `x as i64`: truncate toward zero, saturate, NaN -> 0. */
int64_t fxr_f64_to_i64_sat(double v) {
    if (v != v) return 0;
    if (v >= 9223372036854775807.0) return INT64_MAX;
    if (v <= -9223372036854775808.0) return INT64_MIN;
    return (int64_t)v;
}

/* Rust `x as usize` (saturating at 0 and SIZE_MAX, NaN -> 0). */
static size_t f64_to_usize_sat(double v) {
    if (v != v || v <= 0.0) return 0;
    if (v >= 18446744073709551615.0) return SIZE_MAX;
    return (size_t)v;
}

/* Rust f64::rem_euclid */
static double rem_euclid(double x, double y) {
    double r = fmod(x, y);
    return (r < 0.0) ? r + fabs(y) : r;
}

/* num-complex from_polar(1.0f32, theta) = (1*cos, 1*sin) with f32 libm. */
static void from_polar1_f32(float theta, float out[2]) {
    out[0] = 1.0f * cosf(theta);
    out[1] = 1.0f * sinf(theta);
}

/* num-complex Mul for Complex<f32>: (a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re) */
static inline void cmul_f32(const float a[2], const float b[2], float o[2]) {
    float re = a[0] * b[0] - a[1] * b[1];
    float im = a[0] * b[1] + a[1] * b[0];
    o[0] = re;
    o[1] = im;
}

static inline void cmul_f64(const double a[2], const double b[2], double o[2]) {
    double re = a[0] * b[0] - a[1] * b[1];
    double im = a[0] * b[1] + a[1] * b[0];
    o[0] = re;
    o[1] = im;
}

/* strip leading/trailing ASCII whitespace into buf (Rust str::trim). */
static void trim_copy(const char *s, size_t len, char *buf, size_t cap) {
    size_t b = 0, e = len;
    while (b < e && isspace((unsigned char)s[b])) b++;
    while (e > b && isspace((unsigned char)s[e - 1])) e--;
    size_t n = e - b;
    if (n >= cap) n = cap - 1;
    memcpy(buf, s + b, n);
    buf[n] = 0;
}

/* Rust usize::from_str: optional '+', decimal digits, no overflow. */
static int parse_usize(const char *t, unsigned long long *out) {
    const char *p = t;
    if (*p == '+') p++;
    if (!*p) return -1;
    unsigned long long v = 0;
    for (; *p; ++p) {
        if (*p < '0' || *p > '9') return -1;
        unsigned long long nv = v * 10ull + (unsigned long long)(*p - '0');
        if (nv / 10ull != v) return -1;
        v = nv;
    }
    *out = v;
    return 0;
}

/* Rust f64::from_str (decimal / inf / nan; no hex, no surrounding space). */
static int parse_f64(const char *t, double *out) {
    if (!*t) return -1;
    for (const char *p = t; *p; ++p)
        if (*p == 'x' || *p == 'X' || isspace((unsigned char)*p)) return -1;
    char *end = NULL;
    errno = 0;
    double v = strtod(t, &end);
    if (end == t || *end != 0) return -1;
    *out = v;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* args.rs                                                                  */
/* ------------------------------------------------------------------------ */

/* args.rs:202-210: external list (index 0 = MSB) -> internal[31-i] = ext[i] */
int fxr_parse_shuffle(const char *list, int32_t internal_out[32]) {
    unsigned long long ext[33];
    int count = 0;
    const char *s = list;
    for (;;) {
        const char *c = strchr(s, ',');
        size_t len = c ? (size_t)(c - s) : strlen(s);
        char tok[64];
        trim_copy(s, len, tok, sizeof tok);
        unsigned long long v;
        if (parse_usize(tok, &v) != 0) return -1;
        if (count >= 33) return -1;
        ext[count++] = v;
        if (!c) break;
        s = c + 1;
    }
    if (count != 32) return -1;
    int seen[32] = {0};
    for (int i = 0; i < 32; ++i) {
        if (ext[i] >= 32) return -1;
        if (seen[ext[i]]++) return -1;
    }
    for (int i = 0; i < 32; ++i) internal_out[31 - i] = (int32_t)ext[i];
    return 0;
}

/* args.rs:194-200 */
int fxr_parse_levels(int bits, const char *list, double *out, int cap) {
    if (bits <= 0 || bits > 24) return -1;
    int expected = 1 << bits;
    int count = 0;
    const char *s = list;
    for (;;) {
        const char *c = strchr(s, ',');
        size_t len = c ? (size_t)(c - s) : strlen(s);
        char tok[128];
        trim_copy(s, len, tok, sizeof tok);
        double v;
        if (parse_f64(tok, &v) != 0) return -1;
        if (count < cap) out[count] = v;
        count++;
        if (!c) break;
        s = c + 1;
    }
    if (count != expected || count > cap) return -1;
    return count;
}

/* main.rs:378-387 */
double fxr_quantization_mean_power(const double *levels, int n) {
    if (n <= 0) return -1.0;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += levels[i] * levels[i];
    double p = s / (double)n;
    if (!isfinite(p) || p <= 0.0) return -1.0;
    return p;
}

/* ------------------------------------------------------------------------ */
/* utils.rs: decode                                                          */
/* ------------------------------------------------------------------------ */

enum { FK_GENERIC = 0, FK_P1, FK_P2, FK_P4, FK_P8 };

struct fxr_decode_plan {
    int bits;
    int shuffle_kind;
    uint32_t input_shifts[32];
    int fast_kind;
    float *levels_f32;
    int nlevels;
    float *lut_usb;      /* 65536 x 8, only for 2-bit identity */
    float *lut_lsb_flip; /* 65536 x 8 */
};

/* utils.rs:173-213 */
int fxr_build_decode_plan(int bits, const int32_t shuffle_in[32],
                          const double *levels, int nlevels,
                          fxr_decode_plan **out) {
    if (bits <= 0 || bits > 24 || nlevels < (1 << bits)) return -1;
    fxr_decode_plan *p = (fxr_decode_plan *)calloc(1, sizeof *p);
    if (!p) return -1;
    p->bits = bits;
    int ident = 1, pair = 1;
    for (int i = 0; i < 32; ++i) {
        p->input_shifts[i] = (uint32_t)shuffle_in[i];
        if (shuffle_in[i] != i) ident = 0;
        if (shuffle_in[i] != (i ^ 1)) pair = 0;
    }
    p->shuffle_kind = ident ? FXR_SHUFFLE_IDENTITY
                            : (pair ? FXR_SHUFFLE_PAIRSWAP : FXR_SHUFFLE_GENERIC);
    switch (bits) {
    case 1: p->fast_kind = FK_P1; break;
    case 2: p->fast_kind = FK_P2; break;
    case 4: p->fast_kind = FK_P4; break;
    case 8: p->fast_kind = FK_P8; break;
    default: p->fast_kind = FK_GENERIC; break;
    }
    p->nlevels = nlevels;
    p->levels_f32 = (float *)malloc(sizeof(float) * (size_t)nlevels);
    for (int i = 0; i < nlevels; ++i) p->levels_f32[i] = (float)levels[i];
    if (bits == 2 && p->shuffle_kind == FXR_SHUFFLE_IDENTITY) {
        p->lut_usb = (float *)malloc(sizeof(float) * 65536 * 8);
        p->lut_lsb_flip = (float *)malloc(sizeof(float) * 65536 * 8);
        for (uint32_t w = 0; w < 65536; ++w) {
            for (int i = 0; i < 8; ++i) {
                uint32_t code = (w >> (2 * i)) & 3u;
                float val = p->levels_f32[code];
                p->lut_usb[w * 8 + i] = val;
                p->lut_lsb_flip[w * 8 + i] = (i & 1) ? -val : val;
            }
        }
    }
    *out = p;
    return 0;
}

void fxr_free_decode_plan(fxr_decode_plan *p) {
    if (!p) return;
    free(p->levels_f32);
    free(p->lut_usb);
    free(p->lut_lsb_flip);
    free(p);
}

int fxr_decode_plan_shuffle_kind(const fxr_decode_plan *p) { return p->shuffle_kind; }

/* utils.rs:159-171 */
uint32_t fxr_apply_shuffle_word(uint32_t word, const fxr_decode_plan *p) {
    if (p->shuffle_kind == FXR_SHUFFLE_PAIRSWAP) {
        word = ((word & 0xAAAAAAAAu) >> 1) | ((word & 0x55555555u) << 1);
    } else if (p->shuffle_kind == FXR_SHUFFLE_GENERIC) {
        uint32_t sh = 0;
        for (int out_bit = 0; out_bit < 32; ++out_bit)
            sh |= ((word >> p->input_shifts[out_bit]) & 1u) << out_bit;
        word = sh;
    }
    return word;
}

static inline uint32_t le32(const uint8_t *b) {
    return (uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) |
           ((uint32_t)b[3] << 24);
}

/* utils.rs:215-376 */
int fxr_decode_block(const uint8_t *raw, size_t raw_len, size_t samples,
                     const fxr_decode_plan *p, float *output, size_t out_len,
                     int lsb_to_usb) {
    if (out_len < samples) return -1;
    const float *lv = p->levels_f32;
    size_t out_idx = 0;
    int odd = 0;
    size_t nwords = raw_len / 4;
    switch (p->fast_kind) {
    case FK_P1:
    case FK_P4:
    case FK_P8: {
        int b = (p->fast_kind == FK_P1) ? 1 : (p->fast_kind == FK_P4 ? 4 : 8);
        uint32_t mask = (1u << b) - 1u;
        int per = 32 / b;
        for (size_t w = 0; w < nwords; ++w) {
            uint32_t word = fxr_apply_shuffle_word(le32(raw + 4 * w), p);
            for (int j = 0; j < per; ++j) {
                if (out_idx >= samples) break;
                float val = lv[word & mask];
                if (lsb_to_usb && odd) val = -val;
                output[out_idx++] = val;
                odd = !odd;
                word >>= b;
            }
        }
        break;
    }
    case FK_P2:
        if (p->shuffle_kind == FXR_SHUFFLE_IDENTITY) {
            const float *lut = lsb_to_usb ? p->lut_lsb_flip : p->lut_usb;
            size_t npairs = raw_len / 2;
            for (size_t q = 0; q < npairs; ++q) {
                uint32_t w = (uint32_t)raw[2 * q] | ((uint32_t)raw[2 * q + 1] << 8);
                if (out_idx + 8 <= samples) {
                    memcpy(output + out_idx, lut + (size_t)w * 8, 8 * sizeof(float));
                    out_idx += 8;
                } else {
                    for (int j = 0; j < 8; ++j) {
                        if (out_idx >= samples) break;
                        float val = lv[w & 3u];
                        if (lsb_to_usb && odd) val = -val;
                        output[out_idx++] = val;
                        odd = !odd;
                        w >>= 2;
                    }
                }
            }
            for (size_t r = npairs * 2; r < raw_len; ++r) {
                uint32_t pk = raw[r];
                for (int j = 0; j < 4; ++j) {
                    if (out_idx >= samples) break;
                    float val = lv[pk & 3u];
                    if (lsb_to_usb && odd) val = -val;
                    output[out_idx++] = val;
                    odd = !odd;
                    pk >>= 2;
                }
            }
        } else {
            for (size_t w = 0; w < nwords; ++w) {
                uint32_t word = fxr_apply_shuffle_word(le32(raw + 4 * w), p);
                for (int j = 0; j < 16; ++j) {
                    if (out_idx >= samples) break;
                    float val = lv[word & 3u];
                    if (lsb_to_usb && odd) val = -val;
                    output[out_idx++] = val;
                    odd = !odd;
                    word >>= 2;
                }
            }
        }
        break;
    default: { /* FK_GENERIC */
        int bits = p->bits;
        uint64_t code_mask = (1ull << bits) - 1ull;
        uint64_t acc = 0;
        int acc_bits = 0;
        for (size_t w = 0; w < nwords; ++w) {
            uint32_t word = fxr_apply_shuffle_word(le32(raw + 4 * w), p);
            acc |= ((uint64_t)word) << acc_bits;
            acc_bits += 32;
            while (acc_bits >= bits && out_idx < samples) {
                float val = lv[acc & code_mask];
                if (lsb_to_usb && odd) val = -val;
                output[out_idx++] = val;
                odd = !odd;
                acc_bits -= bits;
                acc >>= bits;
            }
        }
        break;
    }
    }
    return (out_idx == samples) ? 0 : -1;
}

/* utils.rs:126-145 */
void fxr_shift_zerofill(float *x, size_t n, int64_t shift) {
    if (n == 0 || shift == 0) return;
    uint64_t s_abs = (shift < 0) ? (uint64_t)(-(shift + 1)) + 1u : (uint64_t)shift;
    if (s_abs >= n) {
        memset(x, 0, n * sizeof *x);
        return;
    }
    size_t s = (size_t)s_abs;
    if (shift > 0) {
        memmove(x + s, x, (n - s) * sizeof *x);
        memset(x, 0, s * sizeof *x);
    } else {
        memmove(x, x + s, (n - s) * sizeof *x);
        memset(x + (n - s), 0, s * sizeof *x);
    }
}

void fxr_shift_zerofill_f64(double *x, size_t n, int64_t shift) {
    if (n == 0 || shift == 0) return;
    uint64_t s_abs = (shift < 0) ? (uint64_t)(-(shift + 1)) + 1u : (uint64_t)shift;
    if (s_abs >= n) {
        memset(x, 0, n * sizeof *x);
        return;
    }
    size_t s = (size_t)s_abs;
    if (shift > 0) {
        memmove(x + s, x, (n - s) * sizeof *x);
        memset(x, 0, s * sizeof *x);
    } else {
        memmove(x, x + s, (n - s) * sizeof *x);
        memset(x + (n - s), 0, s * sizeof *x);
    }
}

/* ------------------------------------------------------------------------ */
/* FFT (restatement of realfft's method: complex FFT of n/2 + post-process) */
/* ------------------------------------------------------------------------ */

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* In-place iterative radix-2 DIT complex FFT, sign -1 (forward) or +1
 * (inverse, unnormalised), f32 data, twiddles from f64. */
static void cfft_f32(float *z, int m, int sign) {
    for (int i = 1, j = 0; i < m; ++i) {
        int bit = m >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            float t0 = z[2 * i], t1 = z[2 * i + 1];
            z[2 * i] = z[2 * j];
            z[2 * i + 1] = z[2 * j + 1];
            z[2 * j] = t0;
            z[2 * j + 1] = t1;
        }
    }
    for (int len = 2; len <= m; len <<= 1) {
        int half = len >> 1;
        for (int k = 0; k < half; ++k) {
            double ang = (double)sign * 2.0 * FXR_PI * (double)k / (double)len;
            float w[2] = {(float)cos(ang), (float)sin(ang)};
            for (int s = 0; s < m; s += len) {
                float *a = z + 2 * (s + k);
                float *b = z + 2 * (s + k + half);
                float t[2];
                cmul_f32(b, w, t);
                float a0 = a[0], a1 = a[1];
                a[0] = a0 + t[0];
                a[1] = a1 + t[1];
                b[0] = a0 - t[0];
                b[1] = a1 - t[1];
            }
        }
    }
}

static void cfft_f64(double *z, int m, int sign) {
    for (int i = 1, j = 0; i < m; ++i) {
        int bit = m >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t0 = z[2 * i], t1 = z[2 * i + 1];
            z[2 * i] = z[2 * j];
            z[2 * i + 1] = z[2 * j + 1];
            z[2 * j] = t0;
            z[2 * j + 1] = t1;
        }
    }
    for (int len = 2; len <= m; len <<= 1) {
        int half = len >> 1;
        for (int k = 0; k < half; ++k) {
            double ang = (double)sign * 2.0 * FXR_PI * (double)k / (double)len;
            double w[2] = {cos(ang), sin(ang)};
            for (int s = 0; s < m; s += len) {
                double *a = z + 2 * (s + k);
                double *b = z + 2 * (s + k + half);
                double t[2];
                cmul_f64(b, w, t);
                double a0 = a[0], a1 = a[1];
                a[0] = a0 + t[0];
                a[1] = a1 + t[1];
                b[0] = a0 - t[0];
                b[1] = a1 - t[1];
            }
        }
    }
}

/* utils.rs:37-59 -> realfft r2c.  z[m] = x[2m] + i x[2m+1]; Z = FFT_M(z);
 * X[k] = E[k] + W_N^k O[k], E = (Z[k]+conj Z[M-k])/2, O = (Z[k]-conj Z[M-k])/(2i). */
int fxr_rfft_f32(int n, const float *in, float *out) {
    if (n < 4 || !is_pow2(n)) return -1;
    int m = n / 2;
    float *z = (float *)malloc(sizeof(float) * 2 * (size_t)m);
    memcpy(z, in, sizeof(float) * (size_t)n);
    cfft_f32(z, m, -1);
    out[0] = z[0] + z[1];
    out[1] = 0.0f;
    out[2 * m] = z[0] - z[1];
    out[2 * m + 1] = 0.0f;
    for (int k = 1; k < m; ++k) {
        float ar = z[2 * k], ai = z[2 * k + 1];
        float br = z[2 * (m - k)], bi = -z[2 * (m - k) + 1];
        float er = 0.5f * (ar + br), ei = 0.5f * (ai + bi);
        float dr = 0.5f * (ar - br), di = 0.5f * (ai - bi);
        double ang = 2.0 * FXR_PI * (double)k / (double)n;
        float c = (float)cos(ang), s = (float)sin(ang);
        out[2 * k] = er + (c * di - s * dr);
        out[2 * k + 1] = ei - (c * dr + s * di);
    }
    free(z);
    return 0;
}

int fxr_rfft_f64(int n, const double *in, double *out) {
    if (n < 4 || !is_pow2(n)) return -1;
    int m = n / 2;
    double *z = (double *)malloc(sizeof(double) * 2 * (size_t)m);
    memcpy(z, in, sizeof(double) * (size_t)n);
    cfft_f64(z, m, -1);
    out[0] = z[0] + z[1];
    out[1] = 0.0;
    out[2 * m] = z[0] - z[1];
    out[2 * m + 1] = 0.0;
    for (int k = 1; k < m; ++k) {
        double ar = z[2 * k], ai = z[2 * k + 1];
        double br = z[2 * (m - k)], bi = -z[2 * (m - k) + 1];
        double er = 0.5 * (ar + br), ei = 0.5 * (ai + bi);
        double dr = 0.5 * (ar - br), di = 0.5 * (ai - bi);
        double ang = 2.0 * FXR_PI * (double)k / (double)n;
        double c = cos(ang), s = sin(ang);
        out[2 * k] = er + (c * di - s * dr);
        out[2 * k + 1] = ei - (c * dr + s * di);
    }
    free(z);
    return 0;
}

/* utils.rs:60-66: c2r (unnormalised, = n*x) followed by *= 1/n in f32.
 * Imaginary parts of DC and Nyquist are ignored. */
int fxr_irfft_f32(int n, const float *in, float *out) {
    if (n < 4 || !is_pow2(n)) return -1;
    int m = n / 2;
    float *z = (float *)malloc(sizeof(float) * 2 * (size_t)m);
    /* k = 0: E' = X0 + XM, O' = X0 - XM (both real) */
    z[0] = in[0] + in[2 * m];
    z[1] = in[0] - in[2 * m];
    for (int k = 1; k < m; ++k) {
        float ar = in[2 * k], ai = in[2 * k + 1];
        float br = in[2 * (m - k)], bi = -in[2 * (m - k) + 1];
        float er = ar + br, ei = ai + bi;
        float dr = ar - br, di = ai - bi;
        double ang = 2.0 * FXR_PI * (double)k / (double)n;
        float c = (float)cos(ang), s = (float)sin(ang);
        /* O' = D * e^{+i ang} */
        float orr = dr * c - di * s;
        float oi = dr * s + di * c;
        /* Z' = E' + i O' */
        z[2 * k] = er - oi;
        z[2 * k + 1] = ei + orr;
    }
    cfft_f32(z, m, +1);
    float scale = 1.0f / (float)n;
    for (int i = 0; i < n; ++i) out[i] = z[i] * scale;
    free(z);
    return 0;
}

/* The same c2r in FP64 (the FP64 oracle's phased .raw branch). */
int fxr_irfft_f64(int n, const double *in, double *out) {
    if (n < 4 || !is_pow2(n)) return -1;
    int m = n / 2;
    double *z = (double *)malloc(sizeof(double) * 2 * (size_t)m);
    z[0] = in[0] + in[2 * m];
    z[1] = in[0] - in[2 * m];
    for (int k = 1; k < m; ++k) {
        double ar = in[2 * k], ai = in[2 * k + 1];
        double br = in[2 * (m - k)], bi = -in[2 * (m - k) + 1];
        double er = ar + br, ei = ai + bi;
        double dr = ar - br, di = ai - bi;
        double ang = 2.0 * FXR_PI * (double)k / (double)n;
        double c = cos(ang), s = sin(ang);
        double orr = dr * c - di * s;
        double oi = dr * s + di * c;
        z[2 * k] = er - oi;
        z[2 * k + 1] = ei + orr;
    }
    cfft_f64(z, m, +1);
    for (int i = 0; i < n; ++i) out[i] = z[i] / (double)n;
    free(z);
    return 0;
}

/* utils.rs:93-124 */
void fxr_apply_delay_and_rate_range(float *spec, size_t len, size_t fft_len,
                                    double df, double delay_s, double rate,
                                    double accel, double t, int is_lsb,
                                    size_t start_bin, size_t end_bin) {
    if (len == 0) return;
    size_t end = end_bin < len ? end_bin : len;
    size_t start = start_bin < end ? start_bin : end;
    if (start >= end) return;
    double total = delay_s + rate * t + 0.5 * accel * (t * t);
    double sign = is_lsb ? 1.0 : -1.0;
    double phase_step = sign * 2.0 * FXR_PI * df * total;
    float rot[2], step[2];
    from_polar1_f32((float)(phase_step * (double)start), rot);
    from_polar1_f32((float)phase_step, step);
    for (size_t b = start; b < end; ++b) {
        float v[2];
        cmul_f32(spec + 2 * b, rot, v);
        spec[2 * b] = v[0];
        spec[2 * b + 1] = v[1];
        float r2[2];
        cmul_f32(rot, step, r2);
        rot[0] = r2[0];
        rot[1] = r2[1];
    }
    if (fft_len % 2 == 0) {
        size_t nyq = fft_len / 2;
        if (nyq >= start && nyq < end && nyq < len) spec[2 * nyq + 1] = 0.0f;
    }
}

/* ------------------------------------------------------------------------ */
/* main.rs: delay model                                                     */
/* ------------------------------------------------------------------------ */

/* main.rs:110-120 */
void fxr_carrier_phase(double f_hz, double tau_s, float out[2]) {
    if (f_hz == 0.0) {
        out[0] = 1.0f;
        out[1] = 0.0f;
        return;
    }
    double two_pi = 2.0 * FXR_PI;
    double phase = -two_pi * f_hz * tau_s;
    double wrapped = rem_euclid(phase + FXR_PI, two_pi) - FXR_PI;
    from_polar1_f32((float)wrapped, out);
}

static void carrier_phase_f64(double f_hz, double tau_s, double out[2]) {
    if (f_hz == 0.0) {
        out[0] = 1.0;
        out[1] = 0.0;
        return;
    }
    double two_pi = 2.0 * FXR_PI;
    double phase = -two_pi * f_hz * tau_s;
    double wrapped = rem_euclid(phase + FXR_PI, two_pi) - FXR_PI;
    out[0] = cos(wrapped);
    out[1] = sin(wrapped);
}

/* main.rs:122-124 */
static double delay_seconds_at_time(double d, double r, double a, double t) {
    return d + r * t + 0.5 * a * (t * t);
}

/* main.rs:126-130 */
void fxr_split_delay(double delay_s, double fs, int64_t *n, double *frac) {
    int64_t ni = fxr_f64_to_i64_sat(round(delay_s * fs));
    *n = ni;
    *frac = delay_s - ((double)ni / fs);
}

/* main.rs:133-149 */
void fxr_xcf_phase_start_step(double df, double frac1, double frac2,
                              size_t sb1, size_t sb2, float start[2],
                              float step[2]) {
    double two_pi_df = -2.0 * FXR_PI * df;
    double step1 = two_pi_df * frac1;
    double step2 = two_pi_df * frac2;
    from_polar1_f32((float)(step1 * (double)sb1 - step2 * (double)sb2), start);
    from_polar1_f32((float)(step1 - step2), step);
}

static void frame_taus(int64_t frame_idx, const fxr_delay_cfg *cfg, double *tau1,
                       double *tau2, double *tf1, double *tf2) {
    double t_mid_local = ((double)frame_idx + 0.5) * cfg->frame_dt;
    double t_mid = t_mid_local + cfg->time_offset_s;
    if (cfg->geom_table && cfg->geom_table_len > 0) {
        double fl = floor(t_mid_local);
        size_t sec_idx = f64_to_usize_sat(fl > 0.0 ? fl : 0.0);
        size_t last = (size_t)cfg->geom_table_len - 1;
        size_t i0 = sec_idx < last ? sec_idx : last;
        size_t i1 = (i0 + 1) < last ? (i0 + 1) : last;
        double frac;
        if (i1 == i0) {
            frac = 0.0;
        } else {
            frac = t_mid_local - (double)i0;
            if (frac < 0.0) frac = 0.0;
            if (frac > 1.0) frac = 1.0;
        }
        double gd_t = cfg->geom_table[i0] * (1.0 - frac) + cfg->geom_table[i1] * frac;
        double net = gd_t + cfg->coarse_delay_s + cfg->delay_user_samples / cfg->fs +
                     cfg->extra_delay_rate_sps * t_mid;
        double c1 = cfg->clock1_delay_s + cfg->clock1_rate_sps * t_mid;
        double c2 = cfg->clock2_delay_s + cfg->clock2_rate_sps * t_mid;
        double rel = c2 - c1;
        double full = net + rel;
        *tau1 = full - cfg->d_seek;
        *tau2 = 0.0;
        *tf1 = full;
        *tf2 = 0.0;
    } else {
        double t1 = delay_seconds_at_time(cfg->net_d1_base, cfg->total_rate1_base,
                                          cfg->total_accel1_base, t_mid);
        double t2 = delay_seconds_at_time(cfg->net_d2_base, cfg->total_rate2_base,
                                          cfg->total_accel2_base, t_mid);
        *tau1 = t1;
        *tau2 = t2;
        *tf1 = t1 + cfg->d_seek;
        *tf2 = t2;
    }
}

/* main.rs:184-241 */
void fxr_compute_frame_delay(int64_t frame_idx, const fxr_delay_cfg *cfg,
                             fxr_frame_delay *out) {
    double tau1, tau2, tf1, tf2;
    frame_taus(frame_idx, cfg, &tau1, &tau2, &tf1, &tf2);
    fxr_split_delay(tau1, cfg->fs, &out->int1, &out->frac1);
    fxr_split_delay(tau2, cfg->fs, &out->int2, &out->frac2);
    fxr_carrier_phase(cfg->lo1_hz, tf1, out->fr_lo1);
    fxr_carrier_phase(cfg->lo2_hz, tf2, out->fr_lo2);
}

void fxr_compute_frame_delay_f64(int64_t frame_idx, const fxr_delay_cfg *cfg,
                                 fxr_frame_delay *out, double fr1[2],
                                 double fr2[2]) {
    double tau1, tau2, tf1, tf2;
    frame_taus(frame_idx, cfg, &tau1, &tau2, &tf1, &tf2);
    fxr_split_delay(tau1, cfg->fs, &out->int1, &out->frac1);
    fxr_split_delay(tau2, cfg->fs, &out->int2, &out->frac2);
    fxr_carrier_phase(cfg->lo1_hz, tf1, out->fr_lo1);
    fxr_carrier_phase(cfg->lo2_hz, tf2, out->fr_lo2);
    carrier_phase_f64(cfg->lo1_hz, tf1, fr1);
    carrier_phase_f64(cfg->lo2_hz, tf2, fr2);
}

/* main.rs:360-372 */
int fxr_band_alignment(int fft, double fs, double c1, double c2, double bw1,
                       double bw2, int64_t out[4]) {
    if (fft <= 0) return -1;
    double df = fs / (double)fft;
    size_t h = (size_t)fft / 2 + 1;
    size_t lim = (fft % 2 == 0) ? h - 1 : h;
    size_t v1 = f64_to_usize_sat(floor((bw1 * 1e6) / df));
    v1 = (v1 == SIZE_MAX ? v1 : v1 + 1);
    if (v1 > lim) v1 = lim;
    size_t v2 = f64_to_usize_sat(floor((bw2 * 1e6) / df));
    v2 = (v2 == SIZE_MAX ? v2 : v2 + 1);
    if (v2 > lim) v2 = lim;
    double sr = (c1 * 1e6 - c2 * 1e6) / df;
    int64_t sb = fxr_f64_to_i64_sat(round(sr));
    int64_t a1s = (-sb > 0) ? -sb : 0;
    int64_t e2 = (int64_t)v2 - sb;
    int64_t a1e = ((int64_t)v1 < e2) ? (int64_t)v1 : e2;
    if (a1e <= a1s) return -1;
    out[0] = sb;
    out[1] = a1s;
    out[2] = a1e;
    out[3] = a1s + sb;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* utils.rs:378-410 quantise                                                 */
/* ------------------------------------------------------------------------ */

static uint32_t shuffle_out_word(uint32_t word, const int32_t *shuffle_out) {
    uint32_t sh = 0;
    for (int pos = 0; pos < 32; ++pos)
        sh |= ((word >> pos) & 1u) << (uint32_t)shuffle_out[pos];
    return sh;
}

long fxr_quantise_frame(const float *x, size_t n, int bits,
                        const double *levels, int nlevels,
                        const int32_t shuffle_out[32], uint8_t *out,
                        size_t cap) {
    if (bits <= 0 || bits > 32) return -1;
    uint64_t code_mask = (bits >= 64) ? ~0ull : ((1ull << bits) - 1ull);
    uint64_t bit_acc = 0;
    int acc_bits = 0;
    size_t w = 0;
    for (size_t i = 0; i < n; ++i) {
        float val = x[i];
        size_t best = 0;
        float min_err = 3.40282347e+38f;
        for (int k = 0; k < nlevels; ++k) {
            float err = fabsf(val - (float)levels[k]);
            if (err < min_err) {
                min_err = err;
                best = (size_t)k;
            }
        }
        bit_acc |= ((uint64_t)best & code_mask) << acc_bits;
        acc_bits += bits;
        while (acc_bits >= 32) {
            uint32_t word = (uint32_t)(bit_acc & 0xFFFFFFFFull);
            uint32_t sh = shuffle_out_word(word, shuffle_out);
            if (w + 4 > cap) return -1;
            out[w] = (uint8_t)sh;
            out[w + 1] = (uint8_t)(sh >> 8);
            out[w + 2] = (uint8_t)(sh >> 16);
            out[w + 3] = (uint8_t)(sh >> 24);
            w += 4;
            bit_acc >>= 32;
            acc_bits -= 32;
        }
    }
    if (acc_bits > 0) {
        uint32_t word = (uint32_t)(bit_acc & 0xFFFFFFFFull);
        uint32_t sh = shuffle_out_word(word, shuffle_out);
        if (w + 4 > cap) return -1;
        out[w] = (uint8_t)sh;
        out[w + 1] = (uint8_t)(sh >> 8);
        out[w + 2] = (uint8_t)(sh >> 16);
        out[w + 3] = (uint8_t)(sh >> 24);
        w += 4;
    }
    return (long)w;
}

/* ------------------------------------------------------------------------ */
/* main.rs:2173-2591 per-sector compute                                      */
/* ------------------------------------------------------------------------ */

typedef struct {
    const fxr_sector_desc *d;
    const uint8_t *raw1, *raw2;
    int64_t first_frame, nf;
    const fxr_delay_cfg *cfg;
    int mode;
    fxr_decode_plan *dp1, *dp2;
    size_t bpf1, bpf2, bpf_o;
    uint8_t *raw_out;
    int64_t frames_per_job;
    int64_t njobs;
    int64_t next_job; /* guarded by mu */
    pthread_mutex_t mu;
    double *part; /* njobs x 5 x half */
    int64_t *fail; /* per job */
} sector_ctx;

enum { P_PH = 0, P_11 = 1, P_22 = 2, P_12R = 3, P_12I = 4, P_N = 5 };

/* One frame of the accumulate-only branch (main.rs:2448-2568), f32 mode. */
static int frame_acc_f32(sector_ctx *c, int64_t i, float *f1, float *f2,
                         float *s1, float *s2, double *acc) {
    const fxr_sector_desc *d = c->d;
    int n = d->fft_len;
    size_t half = (size_t)n / 2 + 1;
    fxr_frame_delay fd;
    fxr_compute_frame_delay(c->first_frame + i, c->cfg, &fd);
    if (fxr_decode_block(c->raw1 + (size_t)i * c->bpf1, c->bpf1, (size_t)n, c->dp1, f1,
                         (size_t)n, d->lsb1) != 0)
        return -1;
    if (fxr_decode_block(c->raw2 + (size_t)i * c->bpf2, c->bpf2, (size_t)n, c->dp2, f2,
                         (size_t)n, d->lsb2) != 0)
        return -1;
    fxr_shift_zerofill(f1, (size_t)n, fd.int1);
    fxr_shift_zerofill(f2, (size_t)n, fd.int2);
    fxr_rfft_f32(n, f1, s1);
    fxr_rfft_f32(n, f2, s2);
    float fr2c[2] = {fd.fr_lo2[0], -fd.fr_lo2[1]};
    float fr_mix[2];
    cmul_f32(fd.fr_lo1, fr2c, fr_mix);
    size_t ov = (size_t)(d->a1e - d->a1s);
    double *a_ph = acc + P_PH * half, *a11 = acc + P_11 * half, *a22 = acc + P_22 * half;
    double *a12r = acc + P_12R * half, *a12i = acc + P_12I * half;
    (void)a_ph;
    int grid1 = (d->out_grid == FXR_GRID_ANT1);
    if (d->need_acf) {
        if (d->acf_overlap_only) {
            for (size_t k = 0; k < ov; ++k) {
                size_t i1 = (size_t)d->a1s + k, i2 = (size_t)d->a2s + k;
                size_t g = grid1 ? i1 : i2;
                a11[g] += (double)(s1[2 * i1] * s1[2 * i1] + s1[2 * i1 + 1] * s1[2 * i1 + 1]);
                a22[g] += (double)(s2[2 * i2] * s2[2 * i2] + s2[2 * i2 + 1] * s2[2 * i2 + 1]);
            }
        } else if (grid1) {
            for (size_t k = 0; k < half; ++k)
                a11[k] += (double)(s1[2 * k] * s1[2 * k] + s1[2 * k + 1] * s1[2 * k + 1]);
            for (size_t k = 0; k < ov; ++k) {
                size_t i1 = (size_t)d->a1s + k, i2 = (size_t)d->a2s + k;
                a22[i1] += (double)(s2[2 * i2] * s2[2 * i2] + s2[2 * i2 + 1] * s2[2 * i2 + 1]);
            }
        } else {
            for (size_t k = 0; k < half; ++k)
                a22[k] += (double)(s2[2 * k] * s2[2 * k] + s2[2 * k + 1] * s2[2 * k + 1]);
            for (size_t k = 0; k < ov; ++k) {
                size_t i1 = (size_t)d->a1s + k, i2 = (size_t)d->a2s + k;
                a11[i2] += (double)(s1[2 * i1] * s1[2 * i1] + s1[2 * i1 + 1] * s1[2 * i1 + 1]);
            }
        }
    }
    if (d->need_xcf) {
        float ph[2], step[2];
        fxr_xcf_phase_start_step(d->df_hz, fd.frac1, fd.frac2, (size_t)d->a1s,
                                 (size_t)d->a2s, ph, step);
        for (size_t k = 0; k < ov; ++k) {
            size_t i1 = (size_t)d->a1s + k, i2 = (size_t)d->a2s + k;
            float c2[2] = {s2[2 * i2], -s2[2 * i2 + 1]};
            float t0[2], t1[2], v[2];
            cmul_f32(s1 + 2 * i1, c2, t0);
            cmul_f32(t0, fr_mix, t1);
            cmul_f32(t1, ph, v);
            size_t g = grid1 ? i1 : i2;
            a12r[g] += (double)v[0];
            a12i[g] += (double)v[1];
            float np[2];
            cmul_f32(ph, step, np);
            ph[0] = np[0];
            ph[1] = np[1];
        }
    }
    return 0;
}

/* FP64 oracle for one frame: exact decode, FP64 FFT, per-bin FP64 phases. */
static int frame_acc_f64(sector_ctx *c, int64_t i, float *f1, float *f2, double *x1,
                         double *x2, double *s1, double *s2, double *acc) {
    const fxr_sector_desc *d = c->d;
    int n = d->fft_len;
    size_t half = (size_t)n / 2 + 1;
    fxr_frame_delay fd;
    double fr1[2], fr2[2];
    fxr_compute_frame_delay_f64(c->first_frame + i, c->cfg, &fd, fr1, fr2);
    if (fxr_decode_block(c->raw1 + (size_t)i * c->bpf1, c->bpf1, (size_t)n, c->dp1, f1,
                         (size_t)n, d->lsb1) != 0)
        return -1;
    if (fxr_decode_block(c->raw2 + (size_t)i * c->bpf2, c->bpf2, (size_t)n, c->dp2, f2,
                         (size_t)n, d->lsb2) != 0)
        return -1;
    for (int k = 0; k < n; ++k) {
        x1[k] = (double)f1[k];
        x2[k] = (double)f2[k];
    }
    fxr_shift_zerofill_f64(x1, (size_t)n, fd.int1);
    fxr_shift_zerofill_f64(x2, (size_t)n, fd.int2);
    fxr_rfft_f64(n, x1, s1);
    fxr_rfft_f64(n, x2, s2);
    double fr2c[2] = {fr2[0], -fr2[1]};
    double fr_mix[2];
    cmul_f64(fr1, fr2c, fr_mix);
    size_t ov = (size_t)(d->a1e - d->a1s);
    double *a11 = acc + P_11 * half, *a22 = acc + P_22 * half;
    double *a12r = acc + P_12R * half, *a12i = acc + P_12I * half;
    int grid1 = (d->out_grid == FXR_GRID_ANT1);
    if (d->need_acf) {
        if (d->acf_overlap_only) {
            for (size_t k = 0; k < ov; ++k) {
                size_t i1 = (size_t)d->a1s + k, i2 = (size_t)d->a2s + k;
                size_t g = grid1 ? i1 : i2;
                a11[g] += s1[2 * i1] * s1[2 * i1] + s1[2 * i1 + 1] * s1[2 * i1 + 1];
                a22[g] += s2[2 * i2] * s2[2 * i2] + s2[2 * i2 + 1] * s2[2 * i2 + 1];
            }
        } else if (grid1) {
            for (size_t k = 0; k < half; ++k)
                a11[k] += s1[2 * k] * s1[2 * k] + s1[2 * k + 1] * s1[2 * k + 1];
            for (size_t k = 0; k < ov; ++k) {
                size_t i1 = (size_t)d->a1s + k, i2 = (size_t)d->a2s + k;
                a22[i1] += s2[2 * i2] * s2[2 * i2] + s2[2 * i2 + 1] * s2[2 * i2 + 1];
            }
        } else {
            for (size_t k = 0; k < half; ++k)
                a22[k] += s2[2 * k] * s2[2 * k] + s2[2 * k + 1] * s2[2 * k + 1];
            for (size_t k = 0; k < ov; ++k) {
                size_t i1 = (size_t)d->a1s + k, i2 = (size_t)d->a2s + k;
                a11[i2] += s1[2 * i1] * s1[2 * i1] + s1[2 * i1 + 1] * s1[2 * i1 + 1];
            }
        }
    }
    if (d->need_xcf) {
        double two_pi_df = -2.0 * FXR_PI * d->df_hz;
        for (size_t k = 0; k < ov; ++k) {
            size_t i1 = (size_t)d->a1s + k, i2 = (size_t)d->a2s + k;
            double ang = two_pi_df * (fd.frac1 * (double)i1 - fd.frac2 * (double)i2);
            double ph[2] = {cos(ang), sin(ang)};
            double c2[2] = {s2[2 * i2], -s2[2 * i2 + 1]};
            double t0[2], t1[2], v[2];
            cmul_f64(s1 + 2 * i1, c2, t0);
            cmul_f64(t0, fr_mix, t1);
            cmul_f64(t1, ph, v);
            size_t g = grid1 ? i1 : i2;
            a12r[g] += v[0];
            a12i[g] += v[1];
        }
    }
    return 0;
}

/* One frame of the write_raw / phased branch (main.rs:2222-2365), f32. */
static int frame_phased_f32(sector_ctx *c, int64_t i, float *f1, float *f2, float *s1,
                            float *s2, float *s1c, float *s2a, float *cb, float *out_t,
                            double *acc) {
    const fxr_sector_desc *d = c->d;
    int n = d->fft_len;
    size_t half = (size_t)n / 2 + 1;
    uint8_t *out_f = c->raw_out ? c->raw_out + (size_t)i * c->bpf_o : NULL;
    fxr_frame_delay fd;
    fxr_compute_frame_delay(c->first_frame + i, c->cfg, &fd);
    if (fxr_decode_block(c->raw1 + (size_t)i * c->bpf1, c->bpf1, (size_t)n, c->dp1, f1,
                         (size_t)n, d->lsb1) != 0 ||
        fxr_decode_block(c->raw2 + (size_t)i * c->bpf2, c->bpf2, (size_t)n, c->dp2, f2,
                         (size_t)n, d->lsb2) != 0) {
        if (out_f) memset(out_f, 0, c->bpf_o);
        return -1;
    }
    fxr_shift_zerofill(f1, (size_t)n, fd.int1);
    fxr_shift_zerofill(f2, (size_t)n, fd.int2);
    fxr_rfft_f32(n, f1, s1);
    fxr_rfft_f32(n, f2, s2);
    fxr_apply_delay_and_rate_range(s1, half, (size_t)n, d->df_hz, fd.frac1, 0.0, 0.0, 0.0, 0,
                                   0, half);
    fxr_apply_delay_and_rate_range(s2, half, (size_t)n, d->df_hz, fd.frac2, 0.0, 0.0, 0.0, 0,
                                   0, half);
    size_t ov = (size_t)(d->a1e - d->a1s);
    float w1 = (float)d->w1, w2 = (float)d->w2;
    memset(s2a, 0, sizeof(float) * 2 * half);
    if (d->out_grid == FXR_GRID_ANT1) {
        for (size_t k = 0; k < ov; ++k)
            cmul_f32(s2 + 2 * ((size_t)d->a2s + k), fd.fr_lo2, s2a + 2 * ((size_t)d->a1s + k));
        for (size_t k = 0; k < half; ++k) cmul_f32(s1 + 2 * k, fd.fr_lo1, s1c + 2 * k);
        for (size_t k = 0; k < half; ++k) {
            cb[2 * k] = s1c[2 * k] * w1 + s2a[2 * k] * w2;
            cb[2 * k + 1] = s1c[2 * k + 1] * w1 + s2a[2 * k + 1] * w2;
        }
    } else {
        /* s2a holds s1_aligned, s1c holds s2c */
        for (size_t k = 0; k < ov; ++k)
            cmul_f32(s1 + 2 * ((size_t)d->a1s + k), fd.fr_lo1, s2a + 2 * ((size_t)d->a2s + k));
        for (size_t k = 0; k < half; ++k) cmul_f32(s2 + 2 * k, fd.fr_lo2, s1c + 2 * k);
        for (size_t k = 0; k < half; ++k) {
            cb[2 * k] = s2a[2 * k] * w1 + s1c[2 * k] * w2;
            cb[2 * k + 1] = s2a[2 * k + 1] * w1 + s1c[2 * k + 1] * w2;
        }
    }
    double *a_ph = acc + P_PH * half;
    for (size_t k = 0; k < half; ++k)
        a_ph[k] += (double)(cb[2 * k] * cb[2 * k] + cb[2 * k + 1] * cb[2 * k + 1]);
    if (out_f) {
        cb[1] = 0.0f;
        if (n % 2 == 0) cb[2 * ((size_t)n / 2) + 1] = 0.0f;
        fxr_irfft_f32(n, cb, out_t);
        if (d->output_lsb)
            for (int k = 1; k < n; k += 2) out_t[k] = -out_t[k];
        long wr = fxr_quantise_frame(out_t, (size_t)n, d->bit_out, d->levels1, d->nlevels1,
                                     d->shuffle1, out_f, c->bpf_o);
        if (wr != (long)c->bpf_o) {
            memset(out_f, 0, c->bpf_o);
            return -1;
        }
    }
    return 0;
}

/* The phased branch in FP64 (the chosen form's reference): FP64 FFTs, each
 * bin's fractional-delay rotation exp(-2 pi i df frac b) and the fringe
 * phasors evaluated in FP64 (no f32 recurrence), the weighted sum and |cb|^2
 * in FP64; for .raw output the FP64 inverse FFT, the output LSB flip, and
 * the quantiser (utils.rs:378-410) on the samples rounded to f32. */
static int frame_phased_f64(sector_ctx *c, int64_t i, float *f1, float *f2, double *x1,
                            double *x2, double *s1, double *s2, double *s1c, double *s2a,
                            double *cb, double *out_t, double *acc) {
    const fxr_sector_desc *d = c->d;
    int n = d->fft_len;
    size_t half = (size_t)n / 2 + 1;
    uint8_t *out_f = c->raw_out ? c->raw_out + (size_t)i * c->bpf_o : NULL;
    fxr_frame_delay fd;
    double fr1[2], fr2[2];
    fxr_compute_frame_delay_f64(c->first_frame + i, c->cfg, &fd, fr1, fr2);
    if (fxr_decode_block(c->raw1 + (size_t)i * c->bpf1, c->bpf1, (size_t)n, c->dp1, f1,
                         (size_t)n, d->lsb1) != 0 ||
        fxr_decode_block(c->raw2 + (size_t)i * c->bpf2, c->bpf2, (size_t)n, c->dp2, f2,
                         (size_t)n, d->lsb2) != 0) {
        if (out_f) memset(out_f, 0, c->bpf_o);
        return -1;
    }
    for (int k = 0; k < n; ++k) {
        x1[k] = (double)f1[k];
        x2[k] = (double)f2[k];
    }
    fxr_shift_zerofill_f64(x1, (size_t)n, fd.int1);
    fxr_shift_zerofill_f64(x2, (size_t)n, fd.int2);
    fxr_rfft_f64(n, x1, s1);
    fxr_rfft_f64(n, x2, s2);
    /* utils.rs:93-124 with rate = accel = t = 0, USB sign, from bin 0 */
    double *sp[2] = {s1, s2};
    double fr[2] = {fd.frac1, fd.frac2};
    for (int st = 0; st < 2; ++st) {
        double step = -2.0 * FXR_PI * d->df_hz * fr[st];
        for (size_t b = 0; b < half; ++b) {
            double r[2] = {cos(step * (double)b), sin(step * (double)b)}, v[2];
            cmul_f64(sp[st] + 2 * b, r, v);
            sp[st][2 * b] = v[0];
            sp[st][2 * b + 1] = v[1];
        }
        if (n % 2 == 0) sp[st][2 * ((size_t)n / 2) + 1] = 0.0;
    }
    size_t ov = (size_t)(d->a1e - d->a1s);
    double w1 = d->w1, w2 = d->w2;
    memset(s2a, 0, sizeof(double) * 2 * half);
    if (d->out_grid == FXR_GRID_ANT1) {
        for (size_t k = 0; k < ov; ++k)
            cmul_f64(s2 + 2 * ((size_t)d->a2s + k), fr2, s2a + 2 * ((size_t)d->a1s + k));
        for (size_t k = 0; k < half; ++k) cmul_f64(s1 + 2 * k, fr1, s1c + 2 * k);
        for (size_t k = 0; k < half; ++k) {
            cb[2 * k] = s1c[2 * k] * w1 + s2a[2 * k] * w2;
            cb[2 * k + 1] = s1c[2 * k + 1] * w1 + s2a[2 * k + 1] * w2;
        }
    } else {
        for (size_t k = 0; k < ov; ++k)
            cmul_f64(s1 + 2 * ((size_t)d->a1s + k), fr1, s2a + 2 * ((size_t)d->a2s + k));
        for (size_t k = 0; k < half; ++k) cmul_f64(s2 + 2 * k, fr2, s1c + 2 * k);
        for (size_t k = 0; k < half; ++k) {
            cb[2 * k] = s2a[2 * k] * w1 + s1c[2 * k] * w2;
            cb[2 * k + 1] = s2a[2 * k + 1] * w1 + s1c[2 * k + 1] * w2;
        }
    }
    double *a_ph = acc + P_PH * half;
    for (size_t k = 0; k < half; ++k) a_ph[k] += cb[2 * k] * cb[2 * k] + cb[2 * k + 1] * cb[2 * k + 1];
    if (out_f) {
        cb[1] = 0.0;
        if (n % 2 == 0) cb[2 * ((size_t)n / 2) + 1] = 0.0;
        fxr_irfft_f64(n, cb, out_t);
        if (d->output_lsb)
            for (int k = 1; k < n; k += 2) out_t[k] = -out_t[k];
        for (int k = 0; k < n; ++k) f1[k] = (float)out_t[k];
        long wr = fxr_quantise_frame(f1, (size_t)n, d->bit_out, d->levels1, d->nlevels1,
                                     d->shuffle1, out_f, c->bpf_o);
        if (wr != (long)c->bpf_o) {
            memset(out_f, 0, c->bpf_o);
            return -1;
        }
    }
    return 0;
}

static void *sector_worker(void *arg) {
    sector_ctx *c = (sector_ctx *)arg;
    const fxr_sector_desc *d = c->d;
    int n = d->fft_len;
    size_t half = (size_t)n / 2 + 1;
    float *f1 = (float *)malloc(sizeof(float) * (size_t)n);
    float *f2 = (float *)malloc(sizeof(float) * (size_t)n);
    float *s1 = (float *)malloc(sizeof(float) * 2 * half);
    float *s2 = (float *)malloc(sizeof(float) * 2 * half);
    float *s1c = (float *)malloc(sizeof(float) * 2 * half);
    float *s2a = (float *)malloc(sizeof(float) * 2 * half);
    float *cb = (float *)malloc(sizeof(float) * 2 * half);
    float *out_t = (float *)malloc(sizeof(float) * (size_t)n);
    double *x1 = (double *)malloc(sizeof(double) * (size_t)n);
    double *x2 = (double *)malloc(sizeof(double) * (size_t)n);
    double *d1 = (double *)malloc(sizeof(double) * 2 * half);
    double *d2 = (double *)malloc(sizeof(double) * 2 * half);
    int f64ph = d->write_raw && c->mode == FXR_MODE_F64;
    double *p1c = f64ph ? (double *)malloc(sizeof(double) * 2 * half) : NULL;
    double *p2a = f64ph ? (double *)malloc(sizeof(double) * 2 * half) : NULL;
    double *pcb = f64ph ? (double *)malloc(sizeof(double) * 2 * half) : NULL;
    double *pout = f64ph ? (double *)malloc(sizeof(double) * (size_t)n) : NULL;
    for (;;) {
        pthread_mutex_lock(&c->mu);
        int64_t job = c->next_job++;
        pthread_mutex_unlock(&c->mu);
        if (job >= c->njobs) break;
        double *acc = c->part + (size_t)job * P_N * half;
        int64_t start = job * c->frames_per_job;
        int64_t end = start + c->frames_per_job;
        if (end > c->nf) end = c->nf;
        int64_t fails = 0;
        for (int64_t i = start; i < end; ++i) {
            int rc;
            if (d->write_raw && c->mode == FXR_MODE_F64)
                rc = frame_phased_f64(c, i, f1, f2, x1, x2, d1, d2, p1c, p2a, pcb, pout, acc);
            else if (d->write_raw)
                rc = frame_phased_f32(c, i, f1, f2, s1, s2, s1c, s2a, cb, out_t, acc);
            else if (c->mode == FXR_MODE_F64)
                rc = frame_acc_f64(c, i, f1, f2, x1, x2, d1, d2, acc);
            else
                rc = frame_acc_f32(c, i, f1, f2, s1, s2, acc);
            if (rc != 0) fails++;
        }
        c->fail[job] = fails;
    }
    free(f1); free(f2); free(s1); free(s2); free(s1c); free(s2a); free(cb); free(out_t);
    free(x1); free(x2); free(d1); free(d2);
    free(p1c); free(p2a); free(pcb); free(pout);
    return NULL;
}

int fxr_sector_run(const fxr_sector_desc *d, const uint8_t *raw1,
                   const uint8_t *raw2, int64_t first_frame, int64_t nf,
                   const fxr_delay_cfg *cfg, int mode, int threads,
                   double *acc_ph, double *acc11, double *acc12,
                   double *acc22, uint8_t *raw_out, int64_t *n_failed) {
    if (!d || d->fft_len < 4 || (d->fft_len & (d->fft_len - 1)) || nf < 0 || threads < 1)
        return -1;
    if (d->a1s < 0 || d->a2s < 0 || d->a1e <= d->a1s) return -1;
    int n = d->fft_len;
    size_t half = (size_t)n / 2 + 1;
    if ((size_t)(d->a2s + (d->a1e - d->a1s)) > half || (size_t)d->a1e > half) return -1;
    sector_ctx c;
    memset(&c, 0, sizeof c);
    c.d = d;
    c.raw1 = raw1;
    c.raw2 = raw2;
    c.first_frame = first_frame;
    c.nf = nf;
    c.cfg = cfg;
    c.mode = mode;
    c.bpf1 = ((size_t)n * (size_t)d->bits1 + 7) / 8;
    c.bpf2 = ((size_t)n * (size_t)d->bits2 + 7) / 8;
    c.bpf_o = ((size_t)n * (size_t)d->bit_out + 7) / 8;
    c.raw_out = d->write_raw ? raw_out : NULL;
    if (fxr_build_decode_plan(d->bits1, d->shuffle1, d->levels1, d->nlevels1, &c.dp1) != 0)
        return -1;
    if (fxr_build_decode_plan(d->bits2, d->shuffle2, d->levels2, d->nlevels2, &c.dp2) != 0) {
        fxr_free_decode_plan(c.dp1);
        return -1;
    }
    /* main.rs:2441 frames_per_job = clamp(nf / (cpu_threads*8), 128, 2048) */
    int64_t ct = d->cpu_threads > 0 ? d->cpu_threads : 1;
    int64_t fpj = nf / (ct * 8 > 0 ? ct * 8 : 1);
    if (fpj < 128) fpj = 128;
    if (fpj > 2048) fpj = 2048;
    c.frames_per_job = fpj;
    c.njobs = (nf + fpj - 1) / fpj;
    c.part = (double *)calloc((size_t)(c.njobs > 0 ? c.njobs : 1) * P_N * half, sizeof(double));
    c.fail = (int64_t *)calloc((size_t)(c.njobs > 0 ? c.njobs : 1), sizeof(int64_t));
    pthread_mutex_init(&c.mu, NULL);
    int nt = threads;
    if (nt > c.njobs) nt = (int)(c.njobs > 0 ? c.njobs : 1);
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nt);
    for (int t = 0; t < nt; ++t) pthread_create(&th[t], NULL, sector_worker, &c);
    for (int t = 0; t < nt; ++t) pthread_join(th[t], NULL);
    free(th);
    pthread_mutex_destroy(&c.mu);
    /* reduce partials in job order (main.rs:2572-2584) */
    double *tot = (double *)calloc(P_N * half, sizeof(double));
    int64_t fails = 0;
    for (int64_t j = 0; j < c.njobs; ++j) {
        const double *p = c.part + (size_t)j * P_N * half;
        for (size_t k = 0; k < P_N * half; ++k) tot[k] += p[k];
        fails += c.fail[j];
    }
    if (acc_ph) memcpy(acc_ph, tot + P_PH * half, sizeof(double) * half);
    if (acc11) memcpy(acc11, tot + P_11 * half, sizeof(double) * half);
    if (acc22) memcpy(acc22, tot + P_22 * half, sizeof(double) * half);
    if (acc12)
        for (size_t k = 0; k < half; ++k) {
            acc12[2 * k] = tot[P_12R * half + k];
            acc12[2 * k + 1] = tot[P_12I * half + k];
        }
    if (n_failed) *n_failed = fails;
    free(tot);
    free(c.part);
    free(c.fail);
    fxr_free_decode_plan(c.dp1);
    fxr_free_decode_plan(c.dp2);
    return 0;
}
