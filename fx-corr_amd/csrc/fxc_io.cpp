// Host-side I/O of the FX path: the packed-sample reader (main.rs:445-516,
// word-aligned seek main.rs:1578-1616), sector normalisation (main.rs:
// 2645-2683) and the .cor writer (cor.rs:10-296).  Plain C++17 + POSIX.
#include "../../include/fxcorr.h"

#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <fcntl.h>
#include <unistd.h>

extern "C" fxc_status fxc_internal_fail(fxc_status s, const char *msg);

namespace {

fxc_status fail(fxc_status s, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return fxc_internal_fail(s, buf);
}

void put_i32(uint8_t *b, size_t off, int32_t v) { memcpy(b + off, &v, 4); }
void put_u32(uint8_t *b, size_t off, uint32_t v) { memcpy(b + off, &v, 4); }
void put_f32(uint8_t *b, size_t off, float v) { memcpy(b + off, &v, 4); }
void put_f64(uint8_t *b, size_t off, double v) { memcpy(b + off, &v, 8); }
void put_i16(uint8_t *b, size_t off, int16_t v) { memcpy(b + off, &v, 2); }
void put_ascii(uint8_t *b, size_t off, size_t len, const char *s) {
    memset(b + off, 0, len);
    if (!s) return;
    const size_t n = strnlen(s, len);
    memcpy(b + off, s, n);
}

// cor.rs:111-143 split_sec_nsec.
fxc_status split_sec_nsec(int64_t unix_sec, double extra, uint32_t *sec_o, uint32_t *nsec_o) {
    if (!std::isfinite(extra)) return fail(FXC_ERR_INVALID_ARG, "non-finite sector time offset");
    double sec = (double)unix_sec + extra;
    double nsec = (sec - std::trunc(sec)) * 1000000000.0;   // f64::fract
    sec = std::floor(sec);
    if (nsec < 0.0) {
        nsec += 1000000000.0;
        sec -= 1.0;
    } else if (nsec >= 1000000000.0) {
        nsec -= 1000000000.0;
        sec += 1.0;
    }
    // `as i64` saturates (NaN -> 0); round() is half away from zero
    auto to_i64 = [](double v) -> int64_t {
        if (v != v) return 0;
        if (v >= 9223372036854775807.0) return INT64_MAX;
        if (v <= -9223372036854775808.0) return INT64_MIN;
        return (int64_t)v;
    };
    int64_t s64 = to_i64(sec);
    int64_t n64 = to_i64(std::round(nsec));
    if (n64 >= 1000000000) {
        s64 += 1;
        n64 = 0;
    }
    if (s64 < 0 || s64 > (int64_t)UINT32_MAX)
        return fail(FXC_ERR_INVALID_ARG, "sector sec out of u32 range for .cor sector header");
    if (n64 < 0 || n64 > (int64_t)UINT32_MAX)
        return fail(FXC_ERR_INVALID_ARG, "sector nsec out of u32 range for .cor sector header");
    *sec_o = (uint32_t)s64;
    *nsec_o = (uint32_t)n64;
    return FXC_OK;
}

void put_clock(uint8_t *b, size_t off, const fxc_cor_clock_model &m) {
    put_u32(b, off, m.sec);
    put_u32(b, off + 4, m.nsec);
    put_f64(b, off + 8, m.delay);
    put_f64(b, off + 16, m.rate);
    put_f64(b, off + 24, m.acel);
    put_f64(b, off + 32, m.jerk);
    put_f64(b, off + 40, m.snap);
}

}  // namespace

// ============================================================== reader
struct fxc_reader {
    int fd = -1;
    int bit_offset = 0;
    int have_prev = 0;
    uint8_t prev = 0;
    std::vector<uint8_t> next;
};

// read_with_padding (main.rs:445-456): fill `len` bytes, zero past EOF.
static fxc_status read_padded(int fd, uint8_t *buf, size_t len, size_t *got) {
    size_t total = 0;
    while (total < len) {
        const ssize_t r = ::read(fd, buf + total, len - total);
        if (r == 0) {
            memset(buf + total, 0, len - total);
            break;
        }
        if (r < 0) {
            if (errno == EINTR) continue;
            return fail(FXC_ERR_IO, "read failed: %s", strerror(errno));
        }
        total += (size_t)r;
    }
    if (got) *got = total;
    return FXC_OK;
}

extern "C" fxc_status fxc_reader_open(const char *path, uint64_t start_byte, int32_t bit_offset,
                                      fxc_reader **out) {
    if (!path || !out) return fail(FXC_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (bit_offset < 0 || bit_offset >= 8) return fail(FXC_ERR_INVALID_ARG, "bit_offset must be in 0..8");
    const int fd = ::open(path, O_RDONLY | O_CLOEXEC);
    if (fd < 0) return fail(FXC_ERR_IO, "open %s: %s", path, strerror(errno));
#ifdef POSIX_FADV_SEQUENTIAL
    (void)posix_fadvise(fd, 0, 0, POSIX_FADV_SEQUENTIAL);
    (void)posix_fadvise(fd, 0, 0, POSIX_FADV_WILLNEED);
#endif
    if (::lseek(fd, (off_t)start_byte, SEEK_SET) < 0) {
        const int e = errno;
        ::close(fd);
        return fail(FXC_ERR_IO, "seek %s to %llu: %s", path, (unsigned long long)start_byte, strerror(e));
    }
    fxc_reader *r = new fxc_reader();
    r->fd = fd;
    r->bit_offset = bit_offset;
    if (bit_offset != 0) {   // main.rs:470-476
        uint8_t b = 0;
        fxc_status st = read_padded(fd, &b, 1, nullptr);
        if (st != FXC_OK) {
            ::close(fd);
            delete r;
            return st;
        }
        r->prev = b;
        r->have_prev = 1;
    }
    *out = r;
    return FXC_OK;
}

// read_packed_with_padding (main.rs:484-499).
extern "C" fxc_status fxc_reader_read(fxc_reader *r, uint8_t *out, size_t len, size_t *got) {
    if (!r || (!out && len)) return fail(FXC_ERR_INVALID_ARG, "null argument");
    if (r->bit_offset == 0) return read_padded(r->fd, out, len, got);
    r->next.resize(len);
    fxc_status st = read_padded(r->fd, r->next.data(), len, got);
    if (st != FXC_OK) return st;
    const int sh = r->bit_offset;
    uint8_t prev = r->have_prev ? r->prev : 0;
    for (size_t i = 0; i < len; ++i) {
        const uint8_t n = r->next[i];
        out[i] = (uint8_t)((prev >> sh) | (uint8_t)(n << (8 - sh)));
        prev = n;
    }
    r->prev = prev;
    r->have_prev = 1;
    return FXC_OK;
}

extern "C" fxc_status fxc_reader_close(fxc_reader *r) {
    if (!r) return FXC_OK;
    if (r->fd >= 0) ::close(r->fd);
    delete r;
    return FXC_OK;
}

// main.rs:1585-1616: start sample floored to a whole 32-bit word.
extern "C" fxc_status fxc_word_aligned_start(uint64_t start_samples, int32_t bits,
                                             uint64_t *actual_samples, uint64_t *start_byte,
                                             int64_t *residual_samples) {
    if (bits < 1 || bits > 32) return fail(FXC_ERR_INVALID_ARG, "bits %d outside 1..32", bits);
    if (!actual_samples || !start_byte || !residual_samples)
        return fail(FXC_ERR_INVALID_ARG, "null argument");
    const uint64_t spw = 32u / (uint64_t)bits;
    const uint64_t actual = (start_samples / spw) * spw;
    if (actual != 0 && (uint64_t)bits > UINT64_MAX / actual)
        return fail(FXC_ERR_INVALID_ARG, "sample offset overflow while converting to bit offset");
    *actual_samples = actual;
    *start_byte = actual * (uint64_t)bits / 8u;
    *residual_samples = (int64_t)actual - (int64_t)start_samples;
    return FXC_OK;
}

// ======================================================== normalisation
extern "C" fxc_status fxc_normalize_sector(const double *acc, int32_t fft_len, int64_t nf,
                                           double level_power1, double level_power2, double w1,
                                           double w2, float *out_ph, float *out11, float *out22,
                                           float *out12) {
    if (!acc || fft_len <= 0 || (fft_len & 1)) return fail(FXC_ERR_INVALID_ARG, "bad argument");
    const int half = fft_len / 2 + 1;
    const double *ph = acc, *a11 = acc + half, *a22 = acc + 2 * half, *re = acc + 3 * half,
                 *im = acc + 4 * half;
    const double norm_base = (double)nf * ((double)fft_len * (double)fft_len);
    const double inv11 = 1.0 / (0.5 * level_power1 * norm_base);
    const double inv22 = 1.0 / (0.5 * level_power2 * norm_base);
    const double inv12 = 1.0 / (0.5 * std::sqrt(level_power1 * level_power2) * norm_base);
    double pref = w1 * w1 * level_power1 + w2 * w2 * level_power2;
    pref = (pref != pref || pref < 1e-12) ? 1e-12 : pref;   // f64::max(pref, 1e-12): NaN -> 1e-12
    const double invph = 1.0 / (0.5 * pref * norm_base);
    for (int k = 0; k < fft_len / 2; ++k) {
        if (out_ph) {
            out_ph[2 * k] = (float)(ph[k] * invph);
            out_ph[2 * k + 1] = 0.0f;
        }
        if (out11) {
            out11[2 * k] = (float)(a11[k] * inv11);
            out11[2 * k + 1] = 0.0f;
        }
        if (out22) {
            out22[2 * k] = (float)(a22[k] * inv22);
            out22[2 * k + 1] = 0.0f;
        }
        if (out12) {
            out12[2 * k] = (float)(re[k] * inv12);
            out12[2 * k + 1] = (float)(im[k] * inv12);
        }
    }
    return FXC_OK;
}

// ================================================================ .cor
struct fxc_cor_writer {
    FILE *fp = nullptr;
    int32_t fft_half = 0;
    int32_t sectors = 0;
};

extern "C" fxc_status fxc_cor_build_header(const fxc_cor_header *cfg, const fxc_cor_station *st1,
                                           const fxc_cor_station *st2, uint8_t out[256]) {
    if (!cfg || !st1 || !st2 || !out) return fail(FXC_ERR_INVALID_ARG, "null argument");
    if (cfg->fft_point <= 0 || (cfg->fft_point % 2) != 0)
        return fail(FXC_ERR_INVALID_ARG, "cor header fft_point must be positive even integer");
    if (cfg->number_of_sector_hint < 0)
        return fail(FXC_ERR_INVALID_ARG, "cor header number_of_sector_hint must be non-negative");
    if (cfg->clock_reference_unix_sec < 0 || cfg->clock_reference_unix_sec > (int64_t)UINT32_MAX)
        return fail(FXC_ERR_INVALID_ARG, "clock_reference_unix_sec out of u32 range for .cor header");
    const uint32_t cref = (uint32_t)cfg->clock_reference_unix_sec;
    memset(out, 0, 256);
    const uint8_t magic[4] = {0x83, 0xF9, 0xA2, 0x3E};   // cor.rs:166
    memcpy(out, magic, 4);
    put_i32(out, 4, 0x01030000);
    put_i32(out, 8, 1);
    put_i32(out, 12, cfg->sampling_speed_hz);
    put_f64(out, 16, cfg->observing_frequency_hz);
    put_i32(out, 24, cfg->fft_point);
    put_i32(out, 28, cfg->number_of_sector_hint);
    const fxc_cor_station *st[2] = {st1, st2};
    for (int i = 0; i < 2; ++i) {
        const size_t b = 32 + 48 * (size_t)i;
        put_ascii(out, b, 16, st[i]->name);
        put_f64(out, b + 16, st[i]->ecef_m[0]);
        put_f64(out, b + 24, st[i]->ecef_m[1]);
        put_f64(out, b + 32, st[i]->ecef_m[2]);
        // `(code as char).to_string()` truncated to one byte: the code itself
        // below 0x80, else the lead byte of its two-byte UTF-8 form
        const uint8_t c = st[i]->code;
        out[b + 40] = c < 0x80 ? c : (uint8_t)(0xC0 | (c >> 6));
    }
    put_ascii(out, 128, 16, cfg->source_name);
    put_f64(out, 144, cfg->source_ra_rad);
    put_f64(out, 152, cfg->source_dec_rad);
    put_u32(out, 160, cref);
    put_u32(out, 164, 0);
    put_u32(out, 208, cref);
    put_u32(out, 212, 0);
    return FXC_OK;
}

extern "C" fxc_status fxc_cor_build_sector_header(int64_t unix_sec, float integ_s,
                                                  const fxc_cor_sector_model *model,
                                                  uint8_t out[128]) {
    if (!out) return fail(FXC_ERR_INVALID_ARG, "null argument");
    if (unix_sec < INT32_MIN || unix_sec > INT32_MAX)
        return fail(FXC_ERR_INVALID_ARG, "timestamp out of i32 range for .cor sector header");
    memset(out, 0, 128);
    put_i32(out, 0, (int32_t)unix_sec);
    uint32_t s0, n0, s1, n1;
    fxc_status st = split_sec_nsec(unix_sec, 0.0, &s0, &n0);
    if (st != FXC_OK) return st;
    const double integ = (std::isfinite(integ_s) && integ_s > 0.0f) ? (double)integ_s : 1.0;
    st = split_sec_nsec(unix_sec, integ, &s1, &n1);
    if (st != FXC_OK) return st;
    put_u32(out, 4, n0);
    put_u32(out, 8, s1);
    put_u32(out, 12, n1);
    if (model) {
        put_clock(out, 16, model->station1);
        put_clock(out, 64, model->station2);
        put_f32(out, 116, model->amp[0]);
        put_f32(out, 120, model->amp[1]);
        put_i16(out, 124, model->phs[0]);
        put_i16(out, 126, model->phs[1]);
    }
    put_f32(out, 112, integ_s);
    return FXC_OK;
}

extern "C" fxc_status fxc_cor_create(const char *path, const fxc_cor_header *cfg,
                                     const fxc_cor_station *st1, const fxc_cor_station *st2,
                                     fxc_cor_writer **out) {
    if (!path || !out) return fail(FXC_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    uint8_t hdr[256];
    fxc_status st = fxc_cor_build_header(cfg, st1, st2, hdr);
    if (st != FXC_OK) return st;
    FILE *fp = fopen(path, "wb");
    if (!fp) return fail(FXC_ERR_IO, "create %s: %s", path, strerror(errno));
    if (fwrite(hdr, 1, 256, fp) != 256) {
        fclose(fp);
        return fail(FXC_ERR_IO, "write %s failed", path);
    }
    fxc_cor_writer *w = new fxc_cor_writer();
    w->fp = fp;
    w->fft_half = cfg->fft_point / 2;
    *out = w;
    return FXC_OK;
}

extern "C" fxc_status fxc_cor_write_sector(fxc_cor_writer *w, int64_t unix_sec, float integ_s,
                                           const float *spectrum, int64_t n_complex,
                                           const fxc_cor_sector_model *model) {
    if (!w || !w->fp || (!spectrum && n_complex)) return fail(FXC_ERR_INVALID_ARG, "null argument");
    if (n_complex != w->fft_half)
        return fail(FXC_ERR_INVALID_ARG, "sector spectrum length mismatch: expected %d, got %lld",
                    w->fft_half, (long long)n_complex);
    uint8_t hdr[128];
    fxc_status st = fxc_cor_build_sector_header(unix_sec, integ_s, model, hdr);
    if (st != FXC_OK) return st;
    if (fwrite(hdr, 1, 128, w->fp) != 128 ||
        fwrite(spectrum, sizeof(float), 2 * (size_t)n_complex, w->fp) != 2 * (size_t)n_complex)
        return fail(FXC_ERR_IO, "write failed: %s", strerror(errno));
    w->sectors += 1;
    return FXC_OK;
}

// Flushes, stores the sector count at offset 28 (cor.rs:286-296), closes.
extern "C" fxc_status fxc_cor_finalize(fxc_cor_writer *w) {
    if (!w) return FXC_OK;
    fxc_status st = FXC_OK;
    if (w->fp) {
        const int32_t n = w->sectors;
        if (fflush(w->fp) != 0 || fseek(w->fp, 28, SEEK_SET) != 0 || fwrite(&n, 4, 1, w->fp) != 1 ||
            fflush(w->fp) != 0)
            st = fail(FXC_ERR_IO, "finalize failed: %s", strerror(errno));
        if (fclose(w->fp) != 0 && st == FXC_OK) st = fail(FXC_ERR_IO, "close failed");
    }
    delete w;
    return st;
}
