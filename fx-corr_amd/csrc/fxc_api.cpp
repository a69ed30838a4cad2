// C-ABI layer of libfxcorr.so: validation, plans, device tables, launches.
// Mirrors the per-sector seam of M-AKIMOTOO/fx-corr (src/main.rs:2173-2591)
// and the helpers it calls; every entry point cites its reference function.
#include "../../include/fxcorr.h"
#include "fxc_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace fxc;

namespace {

thread_local std::string g_last_error;

fxc_status fail(fxc_status code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

}  // namespace

// Error recording shared with the host I/O unit (fxc_io.cpp).
extern "C" fxc_status fxc_internal_fail(fxc_status s, const char *msg) { return fail(s, "%s", msg); }

namespace {

#define HIP_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(FXC_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));     \
    } while (0)

bool is_pow2(int64_t n) { return n > 0 && (n & (n - 1)) == 0; }

// Device buffer that grows on demand; freed with the plan.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            if (e != hipSuccess) return e;
            p = nullptr;
            cap = 0;
        }
        size_t want = std::max<size_t>(bytes, 256);
        hipError_t e = hipMalloc(&p, want);
        if (e == hipSuccess) cap = want;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Rust usize::from_str / f64::from_str for the list parsers (args.rs).
bool parse_usize_tok(const std::string &t, unsigned long long &out) {
    size_t i = 0;
    if (i < t.size() && t[i] == '+') i++;
    if (i >= t.size()) return false;
    unsigned long long v = 0;
    for (; i < t.size(); ++i) {
        if (t[i] < '0' || t[i] > '9') return false;
        unsigned long long nv = v * 10ull + (unsigned long long)(t[i] - '0');
        if (nv / 10ull != v) return false;
        v = nv;
    }
    out = v;
    return true;
}

bool parse_f64_tok(const std::string &t, double &out) {
    if (t.empty()) return false;
    for (char c : t)
        if (c == 'x' || c == 'X' || isspace((unsigned char)c)) return false;
    char *end = nullptr;
    double v = strtod(t.c_str(), &end);
    if (end == t.c_str() || *end != 0) return false;
    out = v;
    return true;
}

std::string trim(const std::string &s) {
    size_t b = 0, e = s.size();
    while (b < e && isspace((unsigned char)s[b])) b++;
    while (e > b && isspace((unsigned char)s[e - 1])) e--;
    return s.substr(b, e - b);
}

std::vector<std::string> split_commas(const char *list) {
    std::vector<std::string> out;
    std::string cur;
    for (const char *p = list; *p; ++p) {
        if (*p == ',') {
            out.push_back(cur);
            cur.clear();
        } else {
            cur.push_back(*p);
        }
    }
    out.push_back(cur);
    return out;
}

int g_num_cus = -1;
std::mutex g_dev_mu;

int num_cus() {
    std::lock_guard<std::mutex> lk(g_dev_mu);
    if (g_num_cus < 0) {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            g_num_cus = prop.multiProcessorCount;
        else
            g_num_cus = 256;
    }
    return g_num_cus;
}

}  // namespace

struct fxc_plan {
    fxc_plan_desc desc;
    std::vector<double> levels[2];
    int N = 0, M = 0, half = 0;
    int T = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    StationDev st[2];
    int32_t decode_ok[2] = {1, 1};
    // device tables
    DevBuf tw_fft, tw_post, lev_dev[2], out_lev, mf_tab;
    int fft_form = 0;  // 0: FP32 Stockham, 1: FP16/FP32-MFMA four-step (N = 256..2048),
                       // 2: four-step through HBM scratch (N = 32768, 65536)
    bool mf_b1 = false;   // fft_form 0 / 2: mf_tab holds the DFT-32 operand table (scale 1)
                          // and the FFT's first pass runs on FP16 MFMA
    // large-N scratch (fft_form 2): chunk of `chunk` frames
    int M1 = 0, M2 = 0;
    int64_t chunk = 0;       // allocated chunk (frames)
    int64_t chunk_use = 0;   // chunk in use (<= chunk; autotuned)
    int64_t param_cap = 0;   // frames the per-frame parameter buffers hold (>= chunk)
    // autotuned choices of the Stockham form
    int variant = 0;         // Stockham: FFT pass sequence (1: radix-16 passes);
                             // MFMA: kernel organization (1: producer/consumer waves)
    int grid_div = 1;        // workgroups = max_wg / grid_div
    DevBuf tw_l1, tw_l2, tw_lm, d_x[2], d_y, d_z[2], d_params, d_lshifts;
    DevBuf tw_l0;   // W_64 (the fused four-step's rows at N = 16384)
    // scratch
    DevBuf part, red_scratch, table_dev, d_raw[2], d_acc, d_raw_out, d_tmp_in, d_tmp_out,
        d_shifts, d_cb, d_tbuf;
    DevBuf d_chain;          // the reference's f32 phase chains of one slab/chunk of frames
    DevBuf d_fringe;         // one chunk's products for the whole-scan cross spectrum
    DevBuf d_err;            // device error word: bit 0 set by a workgroup whose hand-over wait ran out
    DevBuf d_cacc;           // fused N = 1024 form: per-consumer FP64 totals (kMfConsumerAccBytes per workgroup)
    int max_wg = 0;
    std::vector<double> table_host;  // last uploaded geometric table
    // timing
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending, pool;
    double timed_ms = 0.0;
    int64_t timed_launches = 0;
    // Graphs that captured work of this plan hold its device buffers by
    // address: while any is alive the scratch may not move and the delay
    // table may not change, and fxc_plan_destroy defers the release to the
    // last graph's destruction.
    int live_graphs = 0;
    bool handle_released = false;

    // Every device buffer of the plan (release, byte count).
    template <class Fn> void each_buf(Fn fn) {
        DevBuf *bufs[] = {&tw_fft, &tw_post, &lev_dev[0], &lev_dev[1], &out_lev, &mf_tab,
                          &tw_l1, &tw_l2, &tw_lm, &tw_l0, &d_x[0], &d_x[1], &d_y, &d_z[0], &d_z[1],
                          &d_params, &d_lshifts, &part, &red_scratch, &table_dev, &d_raw[0],
                          &d_raw[1], &d_acc, &d_raw_out, &d_tmp_in, &d_tmp_out, &d_shifts,
                          &d_cb, &d_tbuf, &d_chain, &d_fringe, &d_err, &d_cacc};
        for (DevBuf *b : bufs) fn(*b);
    }
};

// ====================================================================== misc
extern "C" const char *fxc_version(void) { return "fxcorr-mi355x 0.1"; }

extern "C" const char *fxc_strerror(fxc_status s) {
    switch (s) {
    case FXC_OK: return "ok";
    case FXC_ERR_INVALID_ARG: return "invalid argument";
    case FXC_ERR_UNSUPPORTED: return "unsupported configuration";
    case FXC_ERR_HIP: return "HIP runtime error";
    case FXC_ERR_NO_OVERLAP: return "no band overlap";
    case FXC_ERR_PARSE: return "parse error";
    case FXC_ERR_IO: return "I/O error";
    case FXC_ERR_NOMEM: return "out of memory";
    default: return "unknown status";
    }
}

extern "C" const char *fxc_last_error(void) { return g_last_error.c_str(); }

extern "C" fxc_status fxc_device_count(int32_t *n) {
    if (!n) return fail(FXC_ERR_INVALID_ARG, "n is null");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
        *n = 0;
        return fail(FXC_ERR_HIP, "hipGetDeviceCount: %s", hipGetErrorString(e));
    }
    *n = c;
    return FXC_OK;
}

extern "C" fxc_status fxc_device_mem_info(uint64_t *free_bytes, uint64_t *total_bytes) {
    if (!free_bytes || !total_bytes) return fail(FXC_ERR_INVALID_ARG, "null argument");
    size_t f = 0, t = 0;
    HIP_TRY(hipMemGetInfo(&f, &t));
    *free_bytes = f;
    *total_bytes = t;
    return FXC_OK;
}

extern "C" fxc_status fxc_dev_alloc(size_t bytes, void **dptr) {
    if (!dptr) return fail(FXC_ERR_INVALID_ARG, "dptr is null");
    HIP_TRY(hipMalloc(dptr, bytes ? bytes : 1));
    return FXC_OK;
}
extern "C" fxc_status fxc_dev_free(void *dptr) {
    if (dptr) HIP_TRY(hipFree(dptr));
    return FXC_OK;
}
extern "C" fxc_status fxc_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream) {
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return FXC_OK;
}
extern "C" fxc_status fxc_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream) {
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return FXC_OK;
}
extern "C" fxc_status fxc_memset_d(void *dst, int value, size_t bytes, void *stream) {
    HIP_TRY(hipMemsetAsync(dst, value, bytes, (hipStream_t)stream));
    return FXC_OK;
}
extern "C" fxc_status fxc_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream) {
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return FXC_OK;
}
extern "C" fxc_status fxc_stream_create(void **stream) {
    if (!stream) return fail(FXC_ERR_INVALID_ARG, "stream is null");
    hipStream_t s;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return FXC_OK;
}
extern "C" fxc_status fxc_stream_destroy(void *stream) {
    if (stream) HIP_TRY(hipStreamDestroy((hipStream_t)stream));
    return FXC_OK;
}
extern "C" fxc_status fxc_stream_sync(void *stream) {
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return FXC_OK;
}
extern "C" fxc_status fxc_event_create(void **ev) {
    if (!ev) return fail(FXC_ERR_INVALID_ARG, "ev is null");
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    *ev = e;
    return FXC_OK;
}
extern "C" fxc_status fxc_event_destroy(void *ev) {
    if (ev) HIP_TRY(hipEventDestroy((hipEvent_t)ev));
    return FXC_OK;
}
extern "C" fxc_status fxc_event_record(void *ev, void *stream) {
    HIP_TRY(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream));
    return FXC_OK;
}
extern "C" fxc_status fxc_stream_wait_event(void *stream, void *ev) {
    if (!ev) return fail(FXC_ERR_INVALID_ARG, "ev is null");
    HIP_TRY(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)ev, 0));
    return FXC_OK;
}
extern "C" fxc_status fxc_event_elapsed_ms(void *a, void *b, float *ms) {
    if (!ms) return fail(FXC_ERR_INVALID_ARG, "ms is null");
    HIP_TRY(hipEventSynchronize((hipEvent_t)b));
    HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
    return FXC_OK;
}

// ============================================================ host helpers
// args.rs:202-210
extern "C" fxc_status fxc_parse_shuffle(const char *list, int32_t internal_out[32]) {
    if (!list || !internal_out) return fail(FXC_ERR_INVALID_ARG, "null argument");
    std::vector<std::string> toks = split_commas(list);
    std::vector<unsigned long long> ext;
    for (auto &tk : toks) {
        unsigned long long v;
        if (!parse_usize_tok(trim(tk), v))
            return fail(FXC_ERR_PARSE, "invalid shuffle entry '%s'", trim(tk).c_str());
        ext.push_back(v);
    }
    if (ext.size() != 32) return fail(FXC_ERR_PARSE, "Shuffle map must contain exactly 32 entries");
    std::vector<unsigned long long> sorted = ext;
    std::sort(sorted.begin(), sorted.end());
    for (unsigned long long i = 0; i < 32; ++i)
        if (sorted[i] != i) return fail(FXC_ERR_PARSE, "Shuffle map must be a permutation of 0..31");
    for (int i = 0; i < 32; ++i) internal_out[31 - i] = (int32_t)ext[i];
    return FXC_OK;
}

// args.rs:194-200
extern "C" fxc_status fxc_parse_levels(int32_t bits, const char *list, double *out, int32_t cap,
                                       int32_t *n_out) {
    if (!list || !out || !n_out) return fail(FXC_ERR_INVALID_ARG, "null argument");
    if (bits == 0) return fail(FXC_ERR_PARSE, "Bit depth must be at least 1");
    if (bits < 0 || bits > 24) return fail(FXC_ERR_PARSE, "bit depth %d out of range", bits);
    const int64_t expected = int64_t(1) << bits;
    std::vector<std::string> toks = split_commas(list);
    std::vector<double> vals;
    for (auto &tk : toks) {
        double v;
        if (!parse_f64_tok(trim(tk), v))
            return fail(FXC_ERR_PARSE, "invalid level '%s'", trim(tk).c_str());
        vals.push_back(v);
    }
    if ((int64_t)vals.size() != expected)
        return fail(FXC_ERR_PARSE, "Expected %lld levels, received %zu", (long long)expected, vals.size());
    if ((int64_t)cap < expected) return fail(FXC_ERR_INVALID_ARG, "output capacity too small");
    std::copy(vals.begin(), vals.end(), out);
    *n_out = (int32_t)expected;
    return FXC_OK;
}

// main.rs:378-387
extern "C" fxc_status fxc_quantization_mean_power(const double *levels, int32_t n, double *p_out) {
    if (!levels || !p_out) return fail(FXC_ERR_INVALID_ARG, "null argument");
    if (n <= 0) return fail(FXC_ERR_INVALID_ARG, "levels are empty");
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += levels[i] * levels[i];
    double p = s / (double)n;
    if (!std::isfinite(p) || p <= 0.0) return fail(FXC_ERR_INVALID_ARG, "mean-square level is invalid: %g", p);
    *p_out = p;
    return FXC_OK;
}

static size_t f64_to_usize_sat(double v) {
    if (v != v || v <= 0.0) return 0;
    if (v >= 18446744073709551615.0) return SIZE_MAX;
    return (size_t)v;
}
static int64_t f64_to_i64_sat_host(double v) {
    if (v != v) return 0;
    if (v >= 9223372036854775807.0) return INT64_MAX;
    if (v <= -9223372036854775808.0) return INT64_MIN;
    return (int64_t)v;
}

// main.rs:360-372.  The reference's usize cast of a negative right edge
// would wrap; here a non-positive overlap is reported as no overlap.
extern "C" fxc_status fxc_compute_band_alignment(int32_t fft, double fs, double c1, double c2,
                                                 double bw1, double bw2, fxc_band_alignment *out) {
    if (!out || fft <= 0 || !(fs > 0.0)) return fail(FXC_ERR_INVALID_ARG, "bad band alignment args");
    double df = fs / (double)fft;
    size_t h = (size_t)fft / 2 + 1;
    size_t lim = (fft % 2 == 0) ? h - 1 : h;
    size_t v1 = f64_to_usize_sat(std::floor((bw1 * 1e6) / df));
    v1 = std::min(v1 == SIZE_MAX ? v1 : v1 + 1, lim);
    size_t v2 = f64_to_usize_sat(std::floor((bw2 * 1e6) / df));
    v2 = std::min(v2 == SIZE_MAX ? v2 : v2 + 1, lim);
    double sr = (c1 * 1e6 - c2 * 1e6) / df;
    int64_t sb = f64_to_i64_sat_host(std::round(sr));
    int64_t a1s = std::max<int64_t>(0, -sb);
    int64_t a1e = std::min<int64_t>((int64_t)v1, (int64_t)v2 - sb);
    if (a1e <= a1s) return fail(FXC_ERR_NO_OVERLAP, "No band overlap");
    out->shift_bins = sb;
    out->a1s = a1s;
    out->a1e = a1e;
    out->a2s = a1s + sb;
    return FXC_OK;
}

// main.rs:2058-2072
extern "C" fxc_status fxc_sector_partition(int64_t total_f, double frame_sec, int64_t *counts,
                                           int64_t cap, int64_t *n_out) {
    if (!n_out || total_f < 0) return fail(FXC_ERR_INVALID_ARG, "bad partition args");
    double total_duration = (double)total_f * frame_sec;
    double r = std::round(total_duration);
    size_t sector_count = f64_to_usize_sat(r > 1.0 ? r : 1.0);
    size_t tf = (size_t)(total_f > 1 ? total_f : 1);
    if (sector_count > tf) sector_count = tf;
    size_t base = (size_t)total_f / sector_count, extra = (size_t)total_f % sector_count;
    int64_t n = 0;
    for (size_t si = 0; si < sector_count; ++si) {
        int64_t nf = (int64_t)(base + (si < extra ? 1 : 0));
        if (nf > 0) {
            if (counts) {
                if (n >= cap) return fail(FXC_ERR_INVALID_ARG, "counts capacity too small");
                counts[n] = nf;
            }
            n++;
        }
    }
    *n_out = n;
    return FXC_OK;
}

// main.rs:126-130
extern "C" fxc_status fxc_split_delay(double d, double fs, int64_t *n_out, double *frac_out) {
    if (!n_out || !frac_out) return fail(FXC_ERR_INVALID_ARG, "null argument");
    int64_t n = f64_to_i64_sat_host(std::round(d * fs));
    *n_out = n;
    *frac_out = d - ((double)n / fs);
    return FXC_OK;
}

// =================================================================== plans
static fxc_status validate_station(const fxc_station_desc &s, int idx) {
    if (s.bits < 1 || s.bits > kMaxBits)
        return fail(FXC_ERR_UNSUPPORTED, "station %d: bits %d outside 1..%d", idx + 1, s.bits, kMaxBits);
    if (!s.levels || s.nlevels != (1 << s.bits))
        return fail(FXC_ERR_INVALID_ARG, "station %d: expected %d levels, got %d", idx + 1,
                    1 << s.bits, s.nlevels);
    int seen[32] = {0};
    for (int i = 0; i < 32; ++i) {
        if (s.shuffle_internal[i] < 0 || s.shuffle_internal[i] > 31 || seen[s.shuffle_internal[i]]++)
            return fail(FXC_ERR_INVALID_ARG, "station %d: shuffle is not a permutation of 0..31", idx + 1);
    }
    return FXC_OK;
}

// utils.rs:173-213 build_decode_plan (+ the decoded-length rule of
// decode_block_into_with_plan, utils.rs:372-374, evaluated once per plan).
static void build_station(fxc_plan *p, int i) {
    const fxc_station_desc &s = p->desc.st[i];
    StationDev &d = p->st[i];
    memset(&d, 0, sizeof d);
    d.bits = s.bits;
    d.lsb = s.lsb ? 1 : 0;
    bool ident = true, pair = true;
    for (int k = 0; k < 32; ++k) {
        d.shifts[k] = s.shuffle_internal[k];
        if (s.shuffle_internal[k] != k) ident = false;
        if (s.shuffle_internal[k] != (k ^ 1)) pair = false;
    }
    d.shuffle_kind = ident ? SHUF_IDENTITY : (pair ? SHUF_PAIRSWAP : SHUF_GENERIC);
    d.bpf = (int32_t)(((int64_t)p->N * s.bits + 7) / 8);
    d.nwords = (d.bpf + 3) / 4;
    d.aligned = (d.bpf % 4 == 0) ? 1 : 0;
    d.code_mask = (s.bits >= 32) ? 0xFFFFFFFFu : ((1u << s.bits) - 1u);
    for (int k = 0; k < 16 && k < s.nlevels; ++k) d.lev4[k] = (float)s.levels[k];
    int64_t samples;
    if (s.bits == 2 && ident) samples = (int64_t)d.bpf * 8 / 2;       // 16-bit LUT path + tail bytes
    else samples = ((int64_t)(d.bpf / 4) * 32) / s.bits;             // whole 32-bit words only
    d.decode_ok = samples >= p->N ? 1 : 0;
    p->decode_ok[i] = d.decode_ok;
}

static void fill_args(fxc_plan *p, SectorArgs &a) {
    memset(&a, 0, sizeof a);
    a.st[0] = p->st[0];
    a.st[1] = p->st[1];
    a.st[0].levels = (const float *)p->lev_dev[0].p;
    a.st[1].levels = (const float *)p->lev_dev[1].p;
    a.tw_fft = (const float2 *)p->tw_fft.p;
    a.tw_post = (const float2 *)p->tw_post.p;
    a.N = p->N;
    a.M = p->M;
    a.a1s = (int32_t)p->desc.a1s;
    a.a1e = (int32_t)p->desc.a1e;
    a.a2s = (int32_t)p->desc.a2s;
    a.ov = (int32_t)(p->desc.a1e - p->desc.a1s);
    a.grid = p->desc.out_grid;
    a.df = p->desc.df_hz;
    a.w1 = p->desc.w1;
    a.w2 = p->desc.w2;
    a.bit_out = p->desc.bit_out;
    a.output_lsb = p->desc.output_lsb;
    a.out_shuffle_ident = 1;
    for (int k = 0; k < 32; ++k) {
        a.out_shifts[k] = p->desc.st[0].shuffle_internal[k];
        if (a.out_shifts[k] != k) a.out_shuffle_ident = 0;
    }
    a.out_levels = (const float *)p->out_lev.p;
    a.out_nlevels = p->desc.st[0].nlevels;
    a.out_sorted = 1;
    for (int k = 1; k < (int)p->levels[0].size(); ++k)
        if (!((float)p->levels[0][k - 1] <= (float)p->levels[0][k])) a.out_sorted = 0;
    // quantiser search: non-decreasing levels within a quarter step of an
    // even spacing start from a guess (mode 2); the result is the bisection's
    a.out_qmode = a.out_sorted ? 1 : 0;
    a.out_l0 = 0.0f;
    a.out_inv_step = 0.0f;
    {
        const int nl = (int)std::min<size_t>(p->levels[0].size(), 256);
        if (a.out_sorted && nl >= 2) {
            const double l0 = (float)p->levels[0][0], ln = (float)p->levels[0][nl - 1];
            const double d = (ln - l0) / (nl - 1);
            bool even = d > 0.0 && std::isfinite(d);
            for (int k = 0; even && k < nl; ++k)
                even = std::fabs((double)(float)p->levels[0][k] - (l0 + k * d)) <= 0.25 * d;
            if (even) {
                a.out_qmode = 2;
                a.out_l0 = (float)l0;
                a.out_inv_step = (float)(1.0 / d);
            }
        }
    }
    a.bpf_out = (int32_t)(((int64_t)p->N * p->desc.bit_out + 7) / 8);
    uint32_t fl = 0;
    if (p->desc.products & FXC_PRODUCT_ACF) fl |= FLAG_ACF;
    if (p->desc.products & FXC_PRODUCT_XCF) fl |= FLAG_XCF;
    if (p->desc.products & FXC_PRODUCT_PHASED) fl |= FLAG_PHASED;
    if (p->desc.products & FXC_PRODUCT_RAW_OUT) fl |= FLAG_RAW_OUT;
    if (p->desc.products & FXC_PRODUCT_ACF_OVERLAP_ONLY) fl |= FLAG_ACF_OVERLAP;
    a.flags = (int32_t)fl;
    // fused MFMA kernels: their table layout; Stockham / HBM four-step: the
    // DFT-32 operand table when the first pass runs on MFMA, else none
    a.mf_tab = (p->fft_form == 1 || p->mf_b1) ? p->mf_tab.p : nullptr;
    a.variant = p->variant;
    a.err = (int32_t *)p->d_err.p;
    a.cons_acc = (double *)p->d_cacc.p;
}

// IEEE binary16 bits of a float, round to nearest even (normals, subnormals,
// overflow to infinity).
static uint16_t f32_to_f16_bits(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const int32_t exp = (int32_t)((x >> 23) & 0xFF) - 127 + 15;
    uint32_t mant = x & 0x7FFFFFu;
    if (((x >> 23) & 0xFF) == 0xFF) return (uint16_t)(sign | 0x7C00u | (mant ? 0x200u : 0u));
    if (exp >= 31) return (uint16_t)(sign | 0x7C00u);
    if (exp <= 0) {
        if (exp < -10) return (uint16_t)sign;
        mant |= 0x800000u;
        const uint32_t shift = (uint32_t)(14 - exp);
        uint32_t h = mant >> shift;
        const uint32_t rem = mant & ((1u << shift) - 1u), halfway = 1u << (shift - 1);
        if (rem > halfway || (rem == halfway && (h & 1u))) h++;
        return (uint16_t)(sign | h);
    }
    uint32_t h = ((uint32_t)exp << 10) | (mant >> 13);
    const uint32_t rem = mant & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return (uint16_t)(sign | h);
}

static float f16_bits_to_f32(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t exp = (h >> 10) & 0x1Fu, mant = h & 0x3FFu;
    float v;
    if (exp == 0) v = std::ldexp((float)mant, -24);
    else if (exp == 31) v = mant ? NAN : INFINITY;
    else v = std::ldexp((float)(mant | 0x400u), (int)exp - 25);
    uint32_t b;
    memcpy(&b, &v, 4);
    b |= sign;
    memcpy(&v, &b, 4);
    return v;
}

// The MFMA four-step form applies when N = 1024 or 2048, the chosen
// precision is requested, ACF/XCF products are computed on a common band
// start (a1s == a2s, so a bin and its mirror M-k share one lane), every
// level (as the f32 the decode produces, utils.rs:186) is exact in FP16, and
// a station-frame is at most 256 packed words (1/2/4/8-bit at N = 1024,
// 1/2/4-bit at N = 2048).
// Every level of both stations exact in FP16: the decoded samples then feed
// an FP16 MFMA operand without rounding.
static bool levels_fp16_exact(const fxc_plan *p) {
    for (int i = 0; i < 2; ++i)
        for (double lv : p->levels[i]) {
            const float f = (float)lv;
            if (!std::isfinite(f) || f16_bits_to_f32(f32_to_f16_bits(f)) != f) return false;
        }
    return true;
}

static bool mfma_eligible(const fxc_plan *p) {
    if ((p->N != 256 && p->N != 512 && p->N != 1024 && p->N != 2048) ||
        p->desc.precision != FXC_PRECISION_CHOSEN)
        return false;
    // phased plans: the producer/consumer kernel with a phased consumer, N = 256..1024
    if ((p->desc.products & (FXC_PRODUCT_PHASED | FXC_PRODUCT_RAW_OUT)) && p->N == 2048) return false;
    if (p->desc.a1s != p->desc.a2s) return false;
    for (int i = 0; i < 2; ++i) {
        const int b = p->desc.st[i].bits;
        if (b != 1 && b != 2 && b != 4 && b != 8) return false;
        if (p->N * b / 32 > 256) return false;
    }
    return levels_fp16_exact(p);
}

// Chosen form outside the fused MFMA kernels (fft_form 0 with whole waves of
// 32-point DFT rows, N = 1024..16384; fft_form 2, N >= 32768): the FFT's
// first pass runs its DFT-32 on FP16 MFMA when the levels are exact in FP16
// (any band alignment).
static bool first_pass_mfma_eligible(const fxc_plan *p) {
    if (p->desc.precision != FXC_PRECISION_CHOSEN) return false;
    if (!(p->fft_form == 2 || (p->fft_form == 0 && p->N >= 1024 && p->N <= 16384))) return false;
    return levels_fp16_exact(p);
}

// Tables of the MFMA four-step kernel (fxc_mfma.hip), MFT_* layout, all in
// [idx][lane] order (lane l: g = l/16, c = l%16).
//  B1 idx (kap*4 + t)*2 + lohi: half8 slot s of B[K][col] with K = (m1 =
//     16 kap + g + 4 (s/2), e = s%2: re|im of z), col = (k1 = c + 16(t&1),
//     part = t/2: re|im of Y); B = 1/2 * DFT-32 (the realfft split's 1/2),
//     stored as hi + lo, lo = B - hi rounded to FP16 (subnormal for the
//     smaller coefficients: hi + lo is within 2^-25 of B).
//  A2 idx rt*8 + b: DFT-16 coefficient of output (k2 = c, rt: re|im) and
//     input (m2 = 4 g + b/2, b%2: re|im).
//  TW2 idx (ct*4 + i)*2 + part: W512^{m2 k1}, m2 = 4 g + i, k1 = c + 16 ct.
//  TWU idx 2j + part: W1024^p, p = l + 64 j.
// N = 2048 (MFT4_*): the same B1, DLUT and SLUT; no A2;
//  TW2 idx ((rt*2 + h)*4 + i)*2 + part: W1024^{m2 k1}, m2 = 16 rt + 4 g + i,
//     k1 = c + 16 h;
//  TWU idx 2j + part: W2048^p, p = l + 64 j, j < 8.
//  DLUT per station: FP16 samples of each byte value (code = low bits
//     first, odd samples negated for LSB stations); 8-bit keeps an odd-byte
//     copy at +256.  SLUT per station: generic shuffle as 4 byte tables.
// The DFT-32 B operands (scale * DFT-32 as FP16 hi + lo) at tab[o_b1...],
// 16 x 64 half8 in the MFT_B1 layout.
static void put_dft32_b1(std::vector<uint8_t> &tab, size_t o_b1, double scale) {
    auto put16 = [&](size_t off, uint16_t v) { memcpy(&tab[off], &v, 2); };
    for (int kap = 0; kap < 2; ++kap)
        for (int t = 0; t < 4; ++t)
            for (int l = 0; l < 64; ++l)
                for (int s = 0; s < 8; ++s) {
                    const int g = l >> 4, c = l & 15;
                    const int k1 = c + 16 * (t & 1), part = t >> 1;
                    const int m1 = 16 * kap + g + 4 * (s >> 1), e = s & 1;
                    const int r = (m1 * k1) % 32;
                    double cs, sn;
                    if (r % 8 == 0) {   // exact quadrant values
                        const int q = r / 8;
                        cs = (q == 0) ? 1.0 : (q == 2 ? -1.0 : 0.0);
                        sn = (q == 1) ? 1.0 : (q == 3 ? -1.0 : 0.0);
                    } else {
                        cs = std::cos(2.0 * M_PI * r / 32.0);
                        sn = std::sin(2.0 * M_PI * r / 32.0);
                    }
                    const double v = scale * (part == 0 ? (e == 0 ? cs : sn) : (e == 0 ? -sn : cs));
                    const uint16_t hi = f32_to_f16_bits((float)v);
                    const uint16_t lo = f32_to_f16_bits((float)(v - (double)f16_bits_to_f32(hi)));
                    const size_t idx_hi = (size_t)((kap * 4 + t) * 2 + 0), idx_lo = idx_hi + 1;
                    put16(o_b1 + ((idx_hi * 64 + l) * 8 + s) * 2, hi);
                    put16(o_b1 + ((idx_lo * 64 + l) * 8 + s) * 2, lo);
                }
}

static std::vector<uint8_t> mfma_tables(const fxc_plan *p) {
    const bool wide = p->N == 2048;
    const size_t o_b1 = wide ? MFT4_B1 : MFT_B1, o_dlut = wide ? MFT4_DLUT : MFT_DLUT;
    const size_t o_slut = wide ? MFT4_SLUT : MFT_SLUT;
    std::vector<uint8_t> tab(wide ? MFT4_BYTES : MFT_BYTES, 0);
    auto put16 = [&](size_t off, uint16_t v) { memcpy(&tab[off], &v, 2); };
    auto putf = [&](size_t off, float v) { memcpy(&tab[off], &v, 4); };
    auto put32 = [&](size_t off, uint32_t v) { memcpy(&tab[off], &v, 4); };
    put_dft32_b1(tab, o_b1, 0.5);   // the realfft split's 1/2 folded in
    if (wide) {
        for (int rt = 0; rt < 2; ++rt)
            for (int h = 0; h < 2; ++h)
                for (int i = 0; i < 4; ++i)
                    for (int l = 0; l < 64; ++l) {
                        const int m2 = 16 * rt + 4 * (l >> 4) + i, k1 = (l & 15) + 16 * h;
                        const double th = -2.0 * M_PI * (double)((m2 * k1) % 1024) / 1024.0;
                        const size_t idx = (size_t)((rt * 2 + h) * 4 + i) * 2;
                        putf(MFT4_TW2 + (idx * 64 + l) * 4, (float)std::cos(th));
                        putf(MFT4_TW2 + ((idx + 1) * 64 + l) * 4, (float)std::sin(th));
                    }
        for (int j = 0; j < 8; ++j)
            for (int l = 0; l < 64; ++l) {
                const double th = -2.0 * M_PI * (double)(l + 64 * j) / 2048.0;
                putf(MFT4_TWU + ((size_t)(2 * j) * 64 + l) * 4, (float)std::cos(th));
                putf(MFT4_TWU + ((size_t)(2 * j + 1) * 64 + l) * 4, (float)std::sin(th));
            }
    }
    for (int rt = 0; rt < 2 && !wide; ++rt)
        for (int b = 0; b < 8; ++b)
            for (int l = 0; l < 64; ++l) {
                const int k2 = l & 15, m2 = 4 * (l >> 4) + (b >> 1), pin = b & 1;
                const double th = 2.0 * M_PI * (double)((m2 * k2) % 16) / 16.0;
                const double cs = std::cos(th), sn = std::sin(th);
                const double v = rt == 0 ? (pin == 0 ? cs : sn) : (pin == 0 ? -sn : cs);
                putf(MFT_A2 + ((size_t)(rt * 8 + b) * 64 + l) * 4, (float)v);
            }
    // N = 256 / 512: the packed kernel's 16 tile rows are rho = R2 q + m2 for
    // frame q of the slot (R2 = N / 64 rows per frame, M = 32 R2), so the
    // twiddle of row rho is W_M^{(rho % R2) k1}; the untangle twiddles are
    // W_N^p of the frame's pairs
    const int M = p->N / 2, R2 = M / 32;
    for (int ct = 0; ct < 2 && !wide; ++ct)
        for (int i = 0; i < 4; ++i)
            for (int l = 0; l < 64; ++l) {
                const int rho = 4 * (l >> 4) + i, m2 = rho % R2, k1 = (l & 15) + 16 * ct;
                const double th = -2.0 * M_PI * (double)((m2 * k1) % M) / (double)M;
                const size_t idx = (size_t)(ct * 4 + i) * 2;
                putf(MFT_TW2 + (idx * 64 + l) * 4, (float)std::cos(th));
                putf(MFT_TW2 + ((idx + 1) * 64 + l) * 4, (float)std::sin(th));
            }
    for (int j = 0; j < 4 && !wide; ++j)
        for (int l = 0; l < 64; ++l) {
            const double th = -2.0 * M_PI * (double)(l + 64 * j) / (double)p->N;
            putf(MFT_TWU + ((size_t)(2 * j) * 64 + l) * 4, (float)std::cos(th));
            putf(MFT_TWU + ((size_t)(2 * j + 1) * 64 + l) * 4, (float)std::sin(th));
        }
    for (int st = 0; st < 2; ++st) {
        const int bits = p->desc.st[st].bits;
        const bool lsb = p->desc.st[st].lsb != 0;
        const uint32_t mask = (1u << bits) - 1u;
        const size_t base = o_dlut + (size_t)st * MFT_DLUT_HALVES * 2;
        auto lev = [&](uint32_t code, bool neg) {
            uint16_t h = f32_to_f16_bits((float)p->levels[st][code]);
            return (uint16_t)(neg ? (h ^ 0x8000u) : h);
        };
        if (bits == 8) {
            for (int v = 0; v < 256; ++v) {
                put16(base + (size_t)v * 2, lev((uint32_t)v, false));
                put16(base + (size_t)(256 + v) * 2, lev((uint32_t)v, lsb));
            }
        } else {
            const int spb = 8 / bits;   // samples per byte (even)
            for (int v = 0; v < 256; ++v)
                for (int i = 0; i < spb; ++i)
                    put16(base + (size_t)(v * spb + i) * 2,
                          lev(((uint32_t)v >> (bits * i)) & mask, lsb && (i & 1)));
        }
        const size_t sbase = o_slut + (size_t)st * 4 * 256 * 4;
        for (int bp = 0; bp < 4; ++bp)
            for (int v = 0; v < 256; ++v) {
                uint32_t o = 0;
                for (int i = 0; i < 32; ++i) {
                    const int src = p->desc.st[st].shuffle_internal[i];
                    if (src >= 8 * bp && src < 8 * bp + 8 && ((v >> (src - 8 * bp)) & 1)) o |= 1u << i;
                }
                put32(sbase + ((size_t)bp * 256 + v) * 4, o);
            }
    }
    return tab;
}

static bool large_fused_possible(const fxc_plan *p);
static bool phased_fused_possible(const fxc_plan *p);

// The four-step plans' per-frame parameter buffers hold at least 8192
// frames (a few MB), so the fused four-step computes a sector's parameters
// in one launch or a few rather than one per chunk.
constexpr size_t kParamFrames = 8192;
static size_t param_frames(size_t chunk) { return std::max(chunk, kParamFrames); }

// The fused four-step's default chunk: both stations' Y (2 M complex per
// frame) in 128 MiB, so that the second pass reads what the first wrote
// from the 256 MB Infinity Cache rather than from HBM (1024 / 512 / 256
// frames at N = 16384 / 32768 / 65536; measured 5-14 % faster than chunks
// twice as large, and faster than half as large).
static int64_t fused_chunk_default(const fxc_plan *p) {
    const int64_t c = std::max<int64_t>(16, ((int64_t)128 << 20) / (16 * (int64_t)p->M));
    return std::min<int64_t>(p->chunk, c);
}

extern "C" fxc_status fxc_plan_create(const fxc_plan_desc *desc, fxc_plan **out) {
    if (!desc || !out) return fail(FXC_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    const int N = desc->fft_len;
    const bool large = N == 32768 || N == 65536;
    if (!is_pow2(N) || !(sector_supported(N) || large))
        return fail(FXC_ERR_UNSUPPORTED, "fft_len %d: the GPU path supports powers of two 64..65536", N);
    for (int i = 0; i < 2; ++i) {
        fxc_status s = validate_station(desc->st[i], i);
        if (s != FXC_OK) return s;
    }
    const int64_t M = N / 2, half = M + 1;
    if (desc->a1s < 0 || desc->a2s < 0 || desc->a1e <= desc->a1s)
        return fail(FXC_ERR_NO_OVERLAP, "empty band overlap [%lld, %lld)", (long long)desc->a1s,
                    (long long)desc->a1e);
    const int64_t ov = desc->a1e - desc->a1s;
    if (desc->a1e > half || desc->a2s + ov > half)
        return fail(FXC_ERR_INVALID_ARG, "band alignment exceeds the %lld-bin half spectrum", (long long)half);
    if (desc->out_grid != FXC_GRID_ANT1 && desc->out_grid != FXC_GRID_ANT2)
        return fail(FXC_ERR_INVALID_ARG, "out_grid must be 0 or 1");
    if (desc->precision != FXC_PRECISION_PLAIN_FP32 && desc->precision != FXC_PRECISION_CHOSEN)
        return fail(FXC_ERR_INVALID_ARG, "unknown precision %d", desc->precision);
    if (!(desc->df_hz > 0.0)) return fail(FXC_ERR_INVALID_ARG, "df_hz must be > 0");
    const uint32_t known = FXC_PRODUCT_ACF | FXC_PRODUCT_XCF | FXC_PRODUCT_PHASED |
                           FXC_PRODUCT_RAW_OUT | FXC_PRODUCT_ACF_OVERLAP_ONLY;
    if (desc->products & ~known) return fail(FXC_ERR_INVALID_ARG, "unknown product bits");
    if (desc->products & (FXC_PRODUCT_PHASED | FXC_PRODUCT_RAW_OUT)) {
        if (desc->products & (FXC_PRODUCT_ACF | FXC_PRODUCT_XCF))
            return fail(FXC_ERR_INVALID_ARG, "phased products are computed by their own plan");
        if (desc->bit_out < 1 || desc->bit_out > 16)
            return fail(FXC_ERR_UNSUPPORTED, "bit_out %d outside 1..16", desc->bit_out);
        if (((int64_t)N * desc->bit_out) % 32 != 0)
            return fail(FXC_ERR_UNSUPPORTED, "fft_len*bit_out must fill whole 32-bit words");
        if (desc->st[0].nlevels > 256)
            return fail(FXC_ERR_UNSUPPORTED, "phased output supports at most 256 station-1 levels");
    }

    int dev_count = 0;
    hipError_t he = hipGetDeviceCount(&dev_count);
    if (he != hipSuccess || dev_count < 1)
        return fail(FXC_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(he));

    fxc_plan *p = new fxc_plan();
    p->desc = *desc;
    p->N = N;
    p->M = (int)M;
    p->half = (int)half;
    p->T = large ? 512 : sector_threads_per_frame(N);
    (void)hipGetDevice(&p->device);
    for (int i = 0; i < 2; ++i) {
        p->levels[i].assign(desc->st[i].levels, desc->st[i].levels + desc->st[i].nlevels);
        p->desc.st[i].levels = p->levels[i].data();
        build_station(p, i);
    }
    auto cleanup = [&](fxc_status s) {
        fxc_plan_destroy(p);
        return s;
    };
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(FXC_ERR_HIP, "hipStreamCreate failed"));

    // twiddles: W_M^k = exp(-2*pi*i*k/M); post: (cos, sin)(2*pi*k/N)
    std::vector<float> twf(2 * M), twp(2 * half);
    for (int64_t k = 0; k < M; ++k) {
        double ang = -2.0 * M_PI * (double)k / (double)M;
        twf[2 * k] = (float)std::cos(ang);
        twf[2 * k + 1] = (float)std::sin(ang);
    }
    for (int64_t k = 0; k < half; ++k) {
        double ang = 2.0 * M_PI * (double)k / (double)N;
        twp[2 * k] = (float)std::cos(ang);
        twp[2 * k + 1] = (float)std::sin(ang);
    }
    if (p->tw_fft.ensure(twf.size() * 4) != hipSuccess || p->tw_post.ensure(twp.size() * 4) != hipSuccess)
        return cleanup(fail(FXC_ERR_NOMEM, "device allocation failed"));
    if (hipMemcpy(p->tw_fft.p, twf.data(), twf.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->tw_post.p, twp.data(), twp.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(FXC_ERR_HIP, "twiddle upload failed"));
    for (int i = 0; i < 2; ++i) {
        std::vector<float> lv(p->levels[i].begin(), p->levels[i].end());
        if (p->lev_dev[i].ensure(lv.size() * 4) != hipSuccess ||
            hipMemcpy(p->lev_dev[i].p, lv.data(), lv.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(FXC_ERR_HIP, "level upload failed"));
    }
    {
        std::vector<float> lv(p->levels[0].begin(), p->levels[0].end());
        if (p->out_lev.ensure(lv.size() * 4) != hipSuccess ||
            hipMemcpy(p->out_lev.p, lv.data(), lv.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(FXC_ERR_HIP, "level upload failed"));
    }
    const int cus = num_cus();
    if (mfma_eligible(p)) {
        std::vector<uint8_t> tab = mfma_tables(p);
        if (p->mf_tab.ensure(tab.size()) != hipSuccess ||
            hipMemcpy(p->mf_tab.p, tab.data(), tab.size(), hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(FXC_ERR_HIP, "MFMA table upload failed"));
        p->fft_form = 1;
        p->T = 64;   // one wave per frame pair
        p->variant = 1;   // producer/consumer waves (the faster organization on gfx950)
    }
    if (large) {
        // four-step M = M1 x M2 with FFT passes of 128/256 points; scratch for
        // a chunk of frames (decoded samples, Y, Z per station, parameters)
        p->fft_form = 2;
        p->M2 = 128;
        p->M1 = (int)M / p->M2;
        const int64_t per_frame = 2 * (int64_t)N * 4 + 3 * (int64_t)M * 8 + 16 +
                                  (int64_t)frame_param_bytes();
        p->chunk = std::max<int64_t>(16, ((int64_t)768 << 20) / per_frame);
        auto table = [](DevBuf &b, int n) {
            std::vector<float> t(2 * (size_t)n);
            for (int k = 0; k < n; ++k) {
                const double ang = -2.0 * M_PI * (double)k / (double)n;
                t[2 * k] = (float)std::cos(ang);
                t[2 * k + 1] = (float)std::sin(ang);
            }
            return b.ensure(t.size() * 4) == hipSuccess &&
                   hipMemcpy(b.p, t.data(), t.size() * 4, hipMemcpyHostToDevice) == hipSuccess;
        };
        const size_t C = (size_t)p->chunk;
        p->chunk_use = p->chunk;
        if (!table(p->tw_l1, 128) || !table(p->tw_l2, 256) || !table(p->tw_lm, (int)M) ||
            p->d_x[0].ensure(C * N * 4) != hipSuccess || p->d_x[1].ensure(C * N * 4) != hipSuccess ||
            p->d_y.ensure(C * M * 8) != hipSuccess || p->d_z[0].ensure(C * M * 8) != hipSuccess ||
            p->d_z[1].ensure(C * M * 8) != hipSuccess ||
            p->d_params.ensure(param_frames(C) * std::max(frame_param_bytes(), phased_param_bytes())) !=
                hipSuccess ||
            p->d_lshifts.ensure(param_frames(C) * 16) != hipSuccess)
            return cleanup(fail(FXC_ERR_NOMEM, "large-FFT scratch allocation failed"));
        p->param_cap = (int64_t)param_frames(C);
    }
    if (first_pass_mfma_eligible(p)) {
        std::vector<uint8_t> tab(MFT_A2 - MFT_B1, 0);
        put_dft32_b1(tab, 0, 1.0);
        if (p->mf_tab.ensure(tab.size()) != hipSuccess ||
            hipMemcpy(p->mf_tab.p, tab.data(), tab.size(), hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(FXC_ERR_HIP, "MFMA table upload failed"));
        p->mf_b1 = true;
    }
    // grid: plain form about 16 waves per CU worth of workgroups (p->T
    // threads each); MFMA form one 8-wave workgroup per CU; large-N form one
    // partial per frame group.  Each workgroup runs a contiguous range of
    // frames.
    if (large_fused_possible(p) || phased_fused_possible(p)) {   // the fused four-step is the default
        p->variant = 1;
        p->M1 = 128;
        p->M2 = (int)M / 128;
        if (p->fft_form == 2) p->chunk_use = fused_chunk_default(p);
        if (p->fft_form == 0) {
            // N = 16384: Y of both stations for a chunk of frames, the frame
            // (or phased) parameters and shifts, the W_M and W_64 tables
            const size_t pb = std::max(frame_param_bytes(), phased_param_bytes());
            const int64_t per_frame = 2 * (int64_t)M * 8 + 16 + (int64_t)pb;
            p->chunk = std::max<int64_t>(16, ((int64_t)512 << 20) / per_frame);
            const size_t C = (size_t)p->chunk;
            auto table = [](DevBuf &b, int n) {
                std::vector<float> t(2 * (size_t)n);
                for (int k = 0; k < n; ++k) {
                    const double ang = -2.0 * M_PI * (double)k / (double)n;
                    t[2 * k] = (float)std::cos(ang);
                    t[2 * k + 1] = (float)std::sin(ang);
                }
                return b.ensure(t.size() * 4) == hipSuccess &&
                       hipMemcpy(b.p, t.data(), t.size() * 4, hipMemcpyHostToDevice) == hipSuccess;
            };
            p->chunk_use = fused_chunk_default(p);
            if (!table(p->tw_l0, 64) || !table(p->tw_lm, (int)M) ||
                p->d_z[0].ensure(C * M * 8) != hipSuccess || p->d_z[1].ensure(C * M * 8) != hipSuccess ||
                p->d_params.ensure(param_frames(C) * pb) != hipSuccess ||
                p->d_lshifts.ensure(param_frames(C) * 16) != hipSuccess)
                return cleanup(fail(FXC_ERR_NOMEM, "fused four-step scratch allocation failed"));
            p->param_cap = (int64_t)param_frames(C);
        }
    }
    p->max_wg = p->fft_form == 1 ? cus
                : p->fft_form == 2 ? kLargeGroups
                                   : std::max(64, (cus * 1024) / p->T);
    if ((p->fft_form == 1 && p->d_cacc.ensure((size_t)p->max_wg * kMfConsumerAccBytes) != hipSuccess) ||
        p->part.ensure((size_t)p->max_wg * kNumProducts * half * sizeof(double)) != hipSuccess ||
        p->red_scratch.ensure((size_t)64 * kNumProducts * half * sizeof(double)) != hipSuccess ||
        p->d_err.ensure(sizeof(int32_t)) != hipSuccess)
        return cleanup(fail(FXC_ERR_NOMEM, "scratch allocation failed"));
    if (hipMemset(p->d_err.p, 0, sizeof(int32_t)) != hipSuccess)
        return cleanup(fail(FXC_ERR_HIP, "error word initialisation failed"));
    *out = p;
    return FXC_OK;
}

// Plans whose release waits for graphs, and the plans each active capture
// has enqueued work of, keyed by the capture's id so that streams forked
// into a capture count too (collected by enqueue_sector, handed to the
// graph by fxc_graph_end).
static std::mutex g_graph_mu;
static std::vector<std::pair<unsigned long long, fxc_plan *>> g_captured;
static std::vector<std::pair<unsigned long long, hipStream_t>> g_capture_origin;   // id -> begin stream

static void plan_free(fxc_plan *p) {
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    for (auto &e : p->pending) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    for (auto &e : p->pool) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    p->each_buf([](DevBuf &b) { b.release(); });
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

extern "C" fxc_status fxc_plan_destroy(fxc_plan *p) {
    if (!p) return FXC_OK;
    {
        std::lock_guard<std::mutex> lk(g_graph_mu);
        if (p->live_graphs > 0) {   // the last graph that captured it frees it
            p->handle_released = true;
            return FXC_OK;
        }
    }
    plan_free(p);
    return FXC_OK;
}

static bool stream_capturing(hipStream_t s) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return s && hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

// Grows a plan buffer that device-entry work reads or writes.  Refused
// while a live graph holds the old address, and during a capture (the
// captured nodes before it would keep the freed one).
static fxc_status grow(fxc_plan *p, DevBuf &b, size_t bytes, hipStream_t s, const char *what) {
    if (bytes <= b.cap) return FXC_OK;
    if (p->live_graphs > 0)
        return fail(FXC_ERR_INVALID_ARG,
                    "%s: the plan's scratch would grow to %zu bytes, but %d live graph(s) captured "
                    "its current buffers; destroy them first, or make the largest call before capturing",
                    what, bytes, p->live_graphs);
    if (stream_capturing(s))
        return fail(FXC_ERR_INVALID_ARG,
                    "%s: the plan's scratch must grow to %zu bytes during a graph capture; run the "
                    "call once before capturing", what, bytes);
    if (b.ensure(bytes) != hipSuccess)
        return fail(FXC_ERR_NOMEM, "%s: device allocation of %zu bytes failed", what, bytes);
    return FXC_OK;
}
#define GROW(p, buf, bytes, s, what)                                  \
    do {                                                              \
        fxc_status g_ = grow((p), (buf), (bytes), (s), (what));       \
        if (g_ != FXC_OK) return g_;                                  \
    } while (0)

// ================================================================ graphs
struct fxc_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<fxc_plan *> plans;   // plans whose buffers the graph captured
};

// Id of the capture `s` takes part in (0: none).
static unsigned long long capture_id(hipStream_t s) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    unsigned long long id = 0;
    if (s && hipStreamGetCaptureInfo(s, &cs, &id) == hipSuccess && cs == hipStreamCaptureStatusActive)
        return id;
    return 0;
}

static std::vector<fxc_plan *> take_captured(unsigned long long id) {
    std::vector<fxc_plan *> out;
    for (auto it = g_captured.begin(); it != g_captured.end();) {
        if (it->first == id) {
            if (std::find(out.begin(), out.end(), it->second) == out.end()) out.push_back(it->second);
            it = g_captured.erase(it);
        } else {
            ++it;
        }
    }
    return out;
}

static void note_captured(fxc_plan *p, hipStream_t s) {
    const unsigned long long id = capture_id(s);
    std::lock_guard<std::mutex> lk(g_graph_mu);
    for (auto &e : g_captured)
        if (e.first == id && e.second == p) return;
    g_captured.emplace_back(id, p);
}

extern "C" fxc_status fxc_graph_begin(void *stream) {
    if (!stream) return fail(FXC_ERR_INVALID_ARG, "graph capture needs a non-null stream");
    {
        std::lock_guard<std::mutex> lk(g_graph_mu);
        // left over from an abandoned capture begun on this stream
        for (auto it = g_capture_origin.begin(); it != g_capture_origin.end();) {
            if (it->second == (hipStream_t)stream) {
                (void)take_captured(it->first);
                it = g_capture_origin.erase(it);
            } else {
                ++it;
            }
        }
    }
    HIP_TRY(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
    std::lock_guard<std::mutex> lk(g_graph_mu);
    g_capture_origin.emplace_back(capture_id((hipStream_t)stream), (hipStream_t)stream);
    return FXC_OK;
}

extern "C" fxc_status fxc_graph_end(void *stream, fxc_graph **out) {
    if (!stream || !out) return fail(FXC_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    std::vector<fxc_plan *> plans;
    {
        const unsigned long long id = capture_id((hipStream_t)stream);
        std::lock_guard<std::mutex> lk(g_graph_mu);
        plans = take_captured(id);
        for (auto it = g_capture_origin.begin(); it != g_capture_origin.end();)
            it = it->first == id ? g_capture_origin.erase(it) : it + 1;
    }
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamEndCapture((hipStream_t)stream, &g));
    hipGraphExec_t e = nullptr;
    hipError_t he = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    if (he != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(FXC_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(he));
    }
    fxc_graph *r = new fxc_graph();
    r->graph = g;
    r->exec = e;
    r->plans = plans;
    {
        std::lock_guard<std::mutex> lk(g_graph_mu);
        for (fxc_plan *p : plans) p->live_graphs++;
    }
    *out = r;
    return FXC_OK;
}

extern "C" fxc_status fxc_graph_launch(fxc_graph *g, void *stream) {
    if (!g) return fail(FXC_ERR_INVALID_ARG, "null graph");
    HIP_TRY(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return FXC_OK;
}

extern "C" fxc_status fxc_graph_destroy(fxc_graph *g) {
    if (!g) return FXC_OK;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    std::vector<fxc_plan *> to_free;
    {
        std::lock_guard<std::mutex> lk(g_graph_mu);
        for (fxc_plan *p : g->plans)
            if (--p->live_graphs == 0 && p->handle_released) to_free.push_back(p);
    }
    for (fxc_plan *p : to_free) plan_free(p);
    delete g;
    return FXC_OK;
}

extern "C" fxc_status fxc_plan_live_graphs(const fxc_plan *p, int32_t *n) {
    if (!p || !n) return fail(FXC_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(g_graph_mu);
    *n = p->live_graphs;
    return FXC_OK;
}

extern "C" fxc_status fxc_plan_device_bytes(const fxc_plan *p, uint64_t *bytes) {
    if (!p || !bytes) return fail(FXC_ERR_INVALID_ARG, "null argument");
    uint64_t t = 0;
    const_cast<fxc_plan *>(p)->each_buf([&t](DevBuf &b) { t += b.cap; });
    *bytes = t;
    return FXC_OK;
}

// Reads and clears the plan's device error word (after the plan's work has
// completed).  A set bit 0 means a workgroup's bounded hand-over wait ran
// out and its partial sums are not valid.
static fxc_status read_error_word(fxc_plan *p, hipStream_t s) {
    if (!p->d_err.p) return FXC_OK;
    int32_t w = 0;
    HIP_TRY(hipMemcpyAsync(&w, p->d_err.p, sizeof w, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (w != 0) {
        HIP_TRY(hipMemsetAsync(p->d_err.p, 0, sizeof w, s));
        HIP_TRY(hipStreamSynchronize(s));
        return fail(FXC_ERR_HIP,
                    "a sector kernel workgroup's bounded hand-over wait ran out (error word 0x%x); "
                    "the spectra of the affected calls are not valid", (unsigned)w);
    }
    return FXC_OK;
}

// The server copies the word behind each job's result (fxc_server.cpp).
extern "C" const int32_t *fxc_internal_plan_error_word(const fxc_plan *p) {
    return (const int32_t *)p->d_err.p;
}

extern "C" fxc_status fxc_plan_check_device_error(fxc_plan *p, void *stream) {
    if (!p) return fail(FXC_ERR_INVALID_ARG, "null plan");
    if (stream_capturing((hipStream_t)stream))
        return fail(FXC_ERR_INVALID_ARG, "the error word cannot be read inside a graph capture");
    return read_error_word(p, (hipStream_t)stream);
}

extern "C" fxc_status fxc_plan_get_info(const fxc_plan *p, fxc_plan_info *info) {
    if (!p || !info) return fail(FXC_ERR_INVALID_ARG, "null argument");
    info->fft_len = p->N;
    info->half = p->half;
    info->bpf1 = p->st[0].bpf;
    info->bpf2 = p->st[1].bpf;
    info->bpf_out = (int32_t)(((int64_t)p->N * p->desc.bit_out + 7) / 8);
    info->decode_ok1 = p->decode_ok[0];
    info->decode_ok2 = p->decode_ok[1];
    info->fft_form = p->fft_form;
    info->threads_per_frame = p->T;
    info->device_id = p->device;
    info->variant = p->variant;
    info->grid_divisor = p->grid_div;
    info->fft_mfma = (p->fft_form == 1 || p->mf_b1) ? 1 : 0;
    return FXC_OK;
}

static void fill_delay(const fxc_delay_cfg &c, const double *dev_table, DelayDev &d) {
    d.frame_dt = c.frame_dt;
    d.fs = c.fs;
    d.time_offset_s = c.time_offset_s;
    d.table = dev_table;
    d.table_len = dev_table ? c.geom_table_len : 0;
    d.coarse = c.coarse_delay_s;
    d.user_samples = c.delay_user_samples;
    d.extra_rate = c.extra_delay_rate_sps;
    d.c1d = c.clock1_delay_s;
    d.c1r = c.clock1_rate_sps;
    d.c2d = c.clock2_delay_s;
    d.c2r = c.clock2_rate_sps;
    d.d_seek = c.d_seek;
    d.nd1 = c.net_d1_base;
    d.tr1 = c.total_rate1_base;
    d.ta1 = c.total_accel1_base;
    d.nd2 = c.net_d2_base;
    d.tr2 = c.total_rate2_base;
    d.ta2 = c.total_accel2_base;
    d.lo1 = c.lo1_hz;
    d.lo2 = c.lo2_hz;
}

// Upload the geometric table when it differs from the last one.
static fxc_status upload_table(fxc_plan *p, const fxc_delay_cfg &c, hipStream_t s,
                               const double **dev_table) {
    *dev_table = nullptr;
    if (!c.geom_table || c.geom_table_len <= 0) return FXC_OK;
    const size_t n = (size_t)c.geom_table_len;
    bool same = p->table_host.size() == n &&
                memcmp(p->table_host.data(), c.geom_table, n * sizeof(double)) == 0;
    if (!same) {
        if (stream_capturing(s))
            return fail(FXC_ERR_INVALID_ARG,
                        "geometric table changed during graph capture: run the call once before capturing");
        if (p->live_graphs > 0)
            return fail(FXC_ERR_INVALID_ARG,
                        "geometric table differs from the one %d live graph(s) of this plan captured; "
                        "destroy them first, or use another plan", p->live_graphs);
        GROW(p, p->table_dev, n * sizeof(double), s, "geometric table");
        p->table_host.assign(c.geom_table, c.geom_table + n);
        HIP_TRY(hipMemcpyAsync(p->table_dev.p, p->table_host.data(), n * sizeof(double),
                               hipMemcpyHostToDevice, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    *dev_table = (const double *)p->table_dev.p;
    return FXC_OK;
}

static fxc_status collect_timing(fxc_plan *p) {
    for (auto &e : p->pending) {
        HIP_TRY(hipEventSynchronize(e.second));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e.first, e.second));
        p->timed_ms += ms;
        p->timed_launches++;
        p->pool.push_back(e);
    }
    p->pending.clear();
    return FXC_OK;
}

// Enqueue the fused kernel + deterministic reduce for one sector.
static int64_t plan_bpf_out(const fxc_plan *p) {
    return ((int64_t)p->N * p->desc.bit_out + 7) / 8;
}

// Phased-array sector (main.rs:2222-2302, 2391-2411): the spectrum kernel over
// chunks of frames that reuse the same per-workgroup partial slots, and per
// chunk, when the .raw output is requested, the inverse FFT (with the LSB
// flip) and the quantiser into d_raw_out.
// The plain form's XCF takes the reference's f32 phase chain (k_phase_chain)
// when the overlap exceeds kChainMinOverlap bins.
static bool xcf_uses_chain(const fxc_plan *p) {
    return p->desc.precision == FXC_PRECISION_PLAIN_FP32 && (p->desc.products & FXC_PRODUCT_XCF) &&
           (p->desc.a1e - p->desc.a1s) > kChainMinOverlap;
}
constexpr size_t kChainSlabBytes = (size_t)256 << 20;   // chain scratch of the fused plain path

// Long phase chains of the chunked paths (phased, large-N plain) are
// computed for a span of whole chunks at a time, up to this much table: each
// chain is one thread's sequential walk, so the frames of one chunk alone (a
// few hundred to a few thousand) would leave most of the GPU idle for the
// length of the walk.  Chains shorter than kChainSpanMinLen stay per chunk:
// their walks are short and a chunk's table stays in the last-level cache
// between the chain and product kernels.
constexpr size_t kChainSpanBytes = (size_t)4 << 30;
constexpr int kChainSpanMinLen = 4096;

// Frames (whole chunks, at least one) whose chains (`len` links, `row_bytes`
// per frame) one launch computes for a sector of nf frames.
static int64_t chain_span(int64_t nf, int64_t chunk, size_t row_bytes, int len) {
    if (len < kChainSpanMinLen) return chunk;
    const int64_t fit = std::max<int64_t>(1, (int64_t)(kChainSpanBytes / row_bytes) / chunk);
    const int64_t need = (nf + chunk - 1) / chunk;
    return std::min(fit, need) * chunk;
}

// Phased sum on the N = 1024 producer/consumer kernel (chosen form): one
// workgroup per CU over a contiguous range of frames; for the .raw output the
// sector runs in chunks of kPhasedMfCbBytes of combined spectra, each
// followed by the inverse FFT and quantiser (no chunks for |cb|^2 alone).
constexpr size_t kPhasedMfCbBytes = (size_t)512 << 20;

static fxc_status enqueue_phased_mf3(fxc_plan *p, const SectorArgs &base, int64_t nf,
                                     uint8_t *d_raw_out, double *d_acc, hipStream_t s) {
    const bool raw = (p->desc.products & FXC_PRODUCT_RAW_OUT) && d_raw_out != nullptr;
    const size_t half = (size_t)p->half;
    const int64_t bpf_o = plan_bpf_out(p);
    // chunks of equal size (a short last chunk would run on a few workgroups)
    const int64_t cmax = raw ? std::max<int64_t>(1, (int64_t)(kPhasedMfCbBytes / (half * 8))) : nf;
    const int64_t nchunks = (nf + cmax - 1) / cmax;
    const int64_t chunk = (nf + nchunks - 1) / nchunks;
    if (raw) GROW(p, p->d_cb, (size_t)chunk * half * 8, s, "phased spectra");
    const int batch = 8;   // whole frames per producer wave
    int64_t fpw = (chunk + p->max_wg - 1) / p->max_wg;
    fpw = ((fpw + batch - 1) / batch) * batch;
    const int n_wg = (int)((chunk + fpw - 1) / fpw);
    for (int64_t c0 = 0; c0 < nf; c0 += chunk) {
        const int64_t cn = std::min(chunk, nf - c0);
        SectorArgs a = base;
        a.raw1 = base.raw1 + c0 * p->st[0].bpf;
        a.raw2 = base.raw2 + c0 * p->st[1].bpf;
        a.first_frame = base.first_frame + c0;
        a.nf = cn;
        a.frames_per_wg = fpw;
        a.part = (double *)p->part.p;
        a.cb_out = raw ? (float2 *)p->d_cb.p : nullptr;
        a.chain = nullptr;
        HIP_TRY(launch_phased_mf3(a, c0 > 0 ? 1 : 0, (int)((cn + fpw - 1) / fpw), s));
        if (raw)
            HIP_TRY(launch_irfft_quantise(a, (const float2 *)p->d_cb.p, cn, d_raw_out + c0 * bpf_o,
                                          p->desc.output_lsb != 0, s));
    }
    HIP_TRY(launch_reduce((const double *)p->part.p, n_wg, p->half, d_acc,
                          (double *)p->red_scratch.p, 64, s));
    return FXC_OK;
}

static fxc_status enqueue_phased(fxc_plan *p, const SectorArgs &base, int64_t nf,
                                 uint8_t *d_raw_out, double *d_acc, hipStream_t s) {
    if (p->fft_form == 1) return enqueue_phased_mf3(p, base, nf, d_raw_out, d_acc, s);
    const bool raw = (p->desc.products & FXC_PRODUCT_RAW_OUT) && d_raw_out != nullptr;
    const size_t half = (size_t)p->half;
    const int64_t bpf_o = plan_bpf_out(p);
    // per-frame scratch: two rotation chains and, for the .raw output, the
    // combined half spectrum (the inverse FFT quantises its time series from
    // LDS; the N * 4 term keeps the chunking of the staged form).  The chunk
    // does not depend on whether .raw is written, so |cb|^2 sums in the same
    // order either way.
    // chosen form: rotations from FP64 phases in the kernel, no chain table
    const int prec = p->desc.precision == FXC_PRECISION_CHOSEN ? 1 : 0;
    const int ld = chain_ld_for((int)half);
    const size_t per_frame = 2 * (size_t)ld * 8 + half * 8 + (size_t)p->N * 4;
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(nf, (int64_t)(((size_t)256 << 20) / per_frame)));
    const int64_t span = chain_span(nf, chunk, 2 * (size_t)ld * 8, (int)half);
    if (!prec) GROW(p, p->d_chain, (size_t)span * 2 * ld * 8, s, "phased rotation chains");
    if (raw) GROW(p, p->d_cb, (size_t)chunk * half * 8, s, "phased spectra");
    int64_t fpw = (chunk + p->max_wg - 1) / p->max_wg;
    fpw = ((fpw + kFrameBatch - 1) / kFrameBatch) * kFrameBatch;
    if (fpw < kFrameBatch) fpw = kFrameBatch;
    const int n_wg = (int)((chunk + fpw - 1) / fpw);
    int64_t span0 = -1;    // first frame (of this sector) of the computed chains
    for (int64_t c0 = 0; c0 < nf; c0 += chunk) {
        const int64_t cn = std::min(chunk, nf - c0);
        SectorArgs a = base;
        a.raw1 = base.raw1 + c0 * p->st[0].bpf;
        a.raw2 = base.raw2 + c0 * p->st[1].bpf;
        a.first_frame = base.first_frame + c0;
        a.nf = cn;
        a.frames_per_wg = fpw;
        a.part = (double *)p->part.p;
        a.cb_out = raw ? (float2 *)p->d_cb.p : nullptr;
        if (!prec) {
            if (span0 < 0 || c0 >= span0 + span) {
                SectorArgs ca = a;
                ca.nf = std::min(span, nf - c0);
                ca.chain_ld = ld;
                HIP_TRY(launch_phase_chain(ca, CHAIN_PHASED, (float2 *)p->d_chain.p, s));
                span0 = c0;
            }
            a.chain = (const float2 *)p->d_chain.p + (size_t)(c0 - span0) * 2 * ld;   // rows 2f + station
            a.chain_ld = ld;
        } else {
            a.chain = nullptr;
        }
        HIP_TRY(launch_phased(a, prec, c0 > 0 ? 1 : 0, (int)((cn + fpw - 1) / fpw), s));
        if (raw) {
            HIP_TRY(launch_irfft_quantise(a, (const float2 *)p->d_cb.p, cn, d_raw_out + c0 * bpf_o,
                                          p->desc.output_lsb != 0, s));
        }
    }
    HIP_TRY(launch_reduce((const double *)p->part.p, n_wg, p->half, d_acc,
                          (double *)p->red_scratch.p, 64, s));
    return FXC_OK;
}

// Phased-array sum at N = 32768 / 65536: the large path's chunks (decode and
// HBM four-step FFT of both stations), the stations' rotation chains over
// whole chunks, the bin-per-thread phased kernel into kLargeGroups partial
// slots and, for the .raw output, the inverse FFT through the same scratch
// and the quantiser.
static fxc_status enqueue_phased_large(fxc_plan *p, const SectorArgs &base, int64_t nf,
                                       uint8_t *d_raw_out, double *d_acc, hipStream_t s) {
    const bool raw = (p->desc.products & FXC_PRODUCT_RAW_OUT) && d_raw_out != nullptr;
    const size_t half = (size_t)p->half;
    const int64_t bpf_o = plan_bpf_out(p);
    const float2 *twl1 = (const float2 *)(p->M1 == 128 ? p->tw_l1.p : p->tw_l2.p);
    const float2 *twl2 = (const float2 *)(p->M2 == 128 ? p->tw_l1.p : p->tw_l2.p);
    const float2 *twm = (const float2 *)p->tw_lm.p;
    const int prec = p->desc.precision == FXC_PRECISION_CHOSEN ? 1 : 0;
    const int ld = chain_ld_for((int)half);
    const int64_t chunk = p->chunk_use;
    const int64_t span = chain_span(nf, chunk, 2 * (size_t)ld * 8, (int)half);
    if (!prec) GROW(p, p->d_chain, (size_t)span * 2 * ld * 8, s, "phased rotation chains");
    if (raw) GROW(p, p->d_cb, (size_t)chunk * half * 8, s, "phased spectra");
    int64_t span0 = -1;    // first frame (of this sector) of the computed chains
    for (int64_t c0 = 0; c0 < nf; c0 += chunk) {
        const int64_t cn = std::min(chunk, nf - c0);
        SectorArgs a = base;
        a.raw1 = base.raw1 + c0 * p->st[0].bpf;
        a.raw2 = base.raw2 + c0 * p->st[1].bpf;
        a.first_frame = base.first_frame + c0;
        a.nf = cn;
        a.part = (double *)p->part.p;
        a.cb_out = raw ? (float2 *)p->d_cb.p : nullptr;
        int64_t *shifts = (int64_t *)p->d_lshifts.p;
        HIP_TRY(launch_phased_params(a, prec, p->d_params.p, shifts, s));
        // chosen: FP16 decode and the MFMA first pass, as the four-step through HBM
        const bool mf = prec && p->mf_b1;
        for (int st = 0; st < 2; ++st) {
            HIP_TRY(launch_decode_frames(a, st, shifts + st, 2, p->d_x[st].p, mf, s));
            HIP_TRY(launch_fft_large(p->d_x[st].p, mf, (float2 *)p->d_y.p,
                                     (float2 *)p->d_z[st].p, cn, p->M1, p->M2, twl1, twl2, twm,
                                     mf ? p->mf_tab.p : nullptr, s));
        }
        if (!prec) {
            if (span0 < 0 || c0 >= span0 + span) {
                SectorArgs ca = a;
                ca.nf = std::min(span, nf - c0);
                ca.chain_ld = ld;
                HIP_TRY(launch_phase_chain(ca, CHAIN_PHASED, (float2 *)p->d_chain.p, s));
                span0 = c0;
            }
            a.chain = (const float2 *)p->d_chain.p + (size_t)(c0 - span0) * 2 * ld;   // rows 2f + station
            a.chain_ld = ld;
        } else {
            a.chain = nullptr;
        }
        HIP_TRY(launch_phased_large(a, prec, (const float2 *)p->d_z[0].p, (const float2 *)p->d_z[1].p,
                                    p->d_params.p, kLargeGroups, c0 > 0 ? 1 : 0, s));
        if (raw) {
            // d_x[0]: conj(Z'), d_z[0]: its FFT, quantised straight from there
            HIP_TRY(launch_irfft_large((const float2 *)p->d_cb.p, cn, p->M1, p->M2, twl1, twl2, twm,
                                       (const float2 *)p->tw_post.p, (float2 *)p->d_x[0].p,
                                       (float2 *)p->d_y.p, (float2 *)p->d_z[0].p,
                                       p->desc.output_lsb != 0, nullptr, s));
            HIP_TRY(launch_quantise_cplx(a, (const float2 *)p->d_z[0].p, cn, p->desc.output_lsb != 0,
                                         d_raw_out + c0 * bpf_o, s));
        }
    }
    HIP_TRY(launch_reduce((const double *)p->part.p, kLargeGroups, p->half, d_acc,
                          (double *)p->red_scratch.p, 64, s));
    return FXC_OK;
}

// N = 32768 / 65536: chunks of frames through HBM scratch; the product
// kernel's kLargeGroups frame-group partial slots accumulate over the chunks
// in order, then one fixed-order reduce.
// The chosen form's fused four-step (decode in the first pass; the second
// pass, untangle and products in one kernel) applies when both stations'
// bins of a product lie on the same rows (band shift a multiple of M1) and
// both code widths are at most 16 bits (two codes per funnel shift).
// The fused four-step runs at M1 = 128 (variant 1 of the large form).
// The fused four-step (M1 = 128): N = 32768 / 65536 (the large form's
// default), and N = 16384 (M2 = 64) as variant 1 of the Stockham form.
static bool large_fused_possible(const fxc_plan *p) {
    const int64_t sb = p->desc.a2s - p->desc.a1s;
    const bool shape = (p->fft_form == 2 && (p->M / 128 == 128 || p->M / 128 == 256)) ||
                       (p->fft_form == 0 && p->N == 16384);
    return shape && p->desc.precision == FXC_PRECISION_CHOSEN && p->mf_b1 &&
           !(p->desc.products & (FXC_PRODUCT_PHASED | FXC_PRODUCT_RAW_OUT)) && (sb % 128) == 0 &&
           p->st[0].bits <= 16 && p->st[1].bits <= 16;
}

static bool large_fused_ok(const fxc_plan *p) {
    return p->variant == 1 && p->M1 == 128 && large_fused_possible(p);
}

// Frames whose parameters one launch computes on the fused four-step: as
// many whole chunks as the parameter buffer (sized for the plan's largest
// chunk) holds, so that smaller chunks do not each pay a launch for them.
static int64_t fused_param_span(const fxc_plan *p) {
    return std::max<int64_t>(1, p->param_cap / p->chunk_use) * p->chunk_use;
}

// The phased sum at N = 16384 on the fused four-step (chosen form): the
// same first pass, the phased bin in the second pass; as for the products,
// both stations' bins of a pair on the same rows (band shift a multiple of
// M1) and code widths of at most 16 bits.
static bool phased_fused_possible(const fxc_plan *p) {
    const int64_t sb = p->desc.a2s - p->desc.a1s;
    return p->fft_form == 0 && p->N == 16384 && p->desc.precision == FXC_PRECISION_CHOSEN &&
           p->mf_b1 && (p->desc.products & (FXC_PRODUCT_PHASED | FXC_PRODUCT_RAW_OUT)) &&
           (sb % 128) == 0 && p->st[0].bits <= 16 && p->st[1].bits <= 16;
}

static bool phased_fused_ok(const fxc_plan *p) {
    return p->variant == 1 && p->M1 == 128 && phased_fused_possible(p);
}

// Frame groups of the fused second pass: enough workgroups (M1 / 2 row
// pairs x groups) for one resident round at the kernel's occupancy (three
// workgroups per CU at M2 = 256, five at M2 = 64 / 128), at most the plan's
// partial slots.
static int large_fused_groups(const fxc_plan *p) {
    const int per_cu = p->M2 == 256 ? 3 : 5;
    return std::max(1, std::min(p->max_wg, num_cus() * per_cu / (p->M1 / 2)));
}

// Phased sum at N = 16384 on the fused four-step: per chunk the phased
// parameters and shifts, the fused first pass of both stations, the second
// pass with the phased bin into the frame-group partial slots (pair-major)
// and, for the .raw output, the inverse real FFT as a four-step too: the
// second pass forms the half-length spectrum of each pair's rows and their
// DFT-64 (rows of 128 x 64 per frame), the column kernel the DFT-128 over
// the rows and the quantiser.
static fxc_status enqueue_phased_fused(fxc_plan *p, const SectorArgs &base, int64_t nf,
                                       uint8_t *d_raw_out, double *d_acc, hipStream_t s) {
    const bool raw = (p->desc.products & FXC_PRODUCT_RAW_OUT) && d_raw_out != nullptr;
    const int64_t bpf_o = plan_bpf_out(p);
    // the .raw second pass holds four waves per SIMD (the power-only one
    // five): one resident round of workgroups at that occupancy
    const int ng = raw ? std::max(1, std::min(p->max_wg, num_cus() * 4 / (p->M1 / 2)))
                       : large_fused_groups(p);
    if (raw) GROW(p, p->d_cb, (size_t)p->chunk_use * p->half * 8, s, "phased spectra");
    const int64_t span = fused_param_span(p);
    int64_t span0 = -1;   // first frame (of this sector) of the computed parameters
    for (int64_t c0 = 0; c0 < nf; c0 += p->chunk_use) {
        const int64_t cn = std::min(p->chunk_use, nf - c0);
        SectorArgs a = base;
        a.raw1 = base.raw1 + c0 * p->st[0].bpf;
        a.raw2 = base.raw2 + c0 * p->st[1].bpf;
        a.first_frame = base.first_frame + c0;
        a.nf = cn;
        a.part = (double *)p->part.p;
        a.cb_out = raw ? (float2 *)p->d_cb.p : nullptr;
        if (span0 < 0 || c0 >= span0 + span) {
            SectorArgs pa = a;
            pa.nf = std::min(span, nf - c0);
            HIP_TRY(launch_phased_params(pa, 1, p->d_params.p, (int64_t *)p->d_lshifts.p, s));
            span0 = c0;
        }
        const uint8_t *pp = (const uint8_t *)p->d_params.p + (size_t)(c0 - span0) * phased_param_bytes();
        const int64_t *shifts = (const int64_t *)p->d_lshifts.p + 2 * (c0 - span0);
        HIP_TRY(launch_large_pass1_fused(a, shifts, (float2 *)p->d_z[0].p, (float2 *)p->d_z[1].p,
                                         p->M1, p->M2, (const float2 *)p->tw_lm.p, s));
        HIP_TRY(launch_large_pass2_phased(a, (const float2 *)p->d_z[0].p, (const float2 *)p->d_z[1].p,
                                          pp, p->M1, p->M2, (const float2 *)p->tw_l0.p,
                                          (const float2 *)p->tw_lm.p, raw, ng, c0 > 0 ? 1 : 0, s));
        if (raw)   // the inverse FFT's row half done: columns and quantiser
            HIP_TRY(launch_irfft_cols_quantise(a, (const float2 *)p->d_cb.p, cn, d_raw_out + c0 * bpf_o,
                                               p->desc.output_lsb != 0, (const float2 *)p->tw_lm.p, s));
    }
    HIP_TRY(launch_reduce_pairs((const double *)p->part.p, ng, p->half, p->M1, p->M2, d_acc,
                                (double *)p->red_scratch.p, 64, s));
    return FXC_OK;
}

static fxc_status enqueue_large(fxc_plan *p, const SectorArgs &base, int64_t nf, double *d_acc,
                                hipStream_t s) {
    const float2 *twl1 = (const float2 *)(p->M1 == 128 ? p->tw_l1.p : p->tw_l2.p);
    const float2 *twl2 = (const float2 *)(p->M2 == 64 ? p->tw_l0.p : p->M2 == 128 ? p->tw_l1.p : p->tw_l2.p);
    if (large_fused_ok(p)) {
        const int ng = large_fused_groups(p);
        const int64_t span = fused_param_span(p);
        int64_t span0 = -1;   // first frame (of this sector) of the computed parameters
        for (int64_t c0 = 0; c0 < nf; c0 += p->chunk_use) {
            const int64_t cn = std::min(p->chunk_use, nf - c0);
            SectorArgs a = base;
            a.raw1 = base.raw1 + c0 * p->st[0].bpf;
            a.raw2 = base.raw2 + c0 * p->st[1].bpf;
            a.first_frame = base.first_frame + c0;
            a.nf = cn;
            a.part = (double *)p->part.p;
            if (span0 < 0 || c0 >= span0 + span) {
                SectorArgs pa = a;
                pa.nf = std::min(span, nf - c0);
                HIP_TRY(launch_frame_params(pa, p->desc.precision, (FrameParam *)p->d_params.p,
                                            (int64_t *)p->d_lshifts.p, s));
                span0 = c0;
            }
            const FrameParam *params = (const FrameParam *)((const uint8_t *)p->d_params.p +
                                                            (size_t)(c0 - span0) * frame_param_bytes());
            const int64_t *shifts = (const int64_t *)p->d_lshifts.p + 2 * (c0 - span0);
            HIP_TRY(launch_large_pass1_fused(a, shifts, (float2 *)p->d_z[0].p, (float2 *)p->d_z[1].p,
                                             p->M1, p->M2, (const float2 *)p->tw_lm.p, s));
            HIP_TRY(launch_large_pass2_mac(a, (const float2 *)p->d_z[0].p, (const float2 *)p->d_z[1].p,
                                           params, p->M1, p->M2, twl2, ng, c0 > 0 ? 1 : 0, s));
        }
        HIP_TRY(launch_reduce_pairs((const double *)p->part.p, ng, p->half, p->M1, p->M2, d_acc,
                                    (double *)p->red_scratch.p, 64, s));
        return FXC_OK;
    }
    const bool chain = xcf_uses_chain(p);
    const int ld = chain ? chain_ld_for(base.ov) : 0;
    int64_t span = 0;      // frames whose chains one launch computes: whole chunks
    if (chain) {
        span = chain_span(nf, p->chunk_use, (size_t)ld * 8, base.ov);
        GROW(p, p->d_chain, (size_t)span * ld * 8, s, "phase chains");
    }
    int64_t span0 = -1;    // first frame (of this sector) of the computed chains
    for (int64_t c0 = 0; c0 < nf; c0 += p->chunk_use) {
        const int64_t cn = std::min(p->chunk_use, nf - c0);
        SectorArgs a = base;
        a.raw1 = base.raw1 + c0 * p->st[0].bpf;
        a.raw2 = base.raw2 + c0 * p->st[1].bpf;
        a.first_frame = base.first_frame + c0;
        a.nf = cn;
        a.part = (double *)p->part.p;
        FrameParam *params = (FrameParam *)p->d_params.p;
        int64_t *shifts = (int64_t *)p->d_lshifts.p;
        HIP_TRY(launch_frame_params(a, p->desc.precision, params, shifts, s));
        for (int st = 0; st < 2; ++st) {
            // chosen form with the MFMA first pass: the decode writes FP16 pairs
            // (exact for these levels), half the bytes of FP32
            HIP_TRY(launch_decode_frames(a, st, shifts + st, 2, p->d_x[st].p, p->mf_b1, s));
            HIP_TRY(launch_fft_large(p->d_x[st].p, p->mf_b1, (float2 *)p->d_y.p,
                                     (float2 *)p->d_z[st].p, cn, p->M1, p->M2, twl1, twl2,
                                     (const float2 *)p->tw_lm.p,
                                     p->mf_b1 ? p->mf_tab.p : nullptr, s));
        }
        if (chain) {
            if (span0 < 0 || c0 >= span0 + span) {
                SectorArgs ca = a;
                ca.nf = std::min(span, nf - c0);
                ca.chain_ld = ld;
                HIP_TRY(launch_phase_chain(ca, CHAIN_XCF, (float2 *)p->d_chain.p, s));
                span0 = c0;
            }
            a.chain = (const float2 *)p->d_chain.p + (size_t)(c0 - span0) * ld;
            a.chain_ld = ld;
        }
        HIP_TRY(launch_mac_large(a, p->desc.precision, (const float2 *)p->d_z[0].p,
                                 (const float2 *)p->d_z[1].p, params, kLargeGroups, c0 > 0 ? 1 : 0,
                                 s));
    }
    HIP_TRY(launch_reduce((const double *)p->part.p, kLargeGroups, p->half, d_acc,
                          (double *)p->red_scratch.p, 64, s));
    return FXC_OK;
}

static fxc_status enqueue_sector(fxc_plan *p, const uint8_t *d_raw1, const uint8_t *d_raw2,
                                 int64_t first_frame, int64_t nf, const fxc_delay_cfg &dc,
                                 double *d_acc, uint8_t *d_raw_out, hipStream_t s) {
    const double *dtab = nullptr;
    fxc_status st = upload_table(p, dc, s, &dtab);
    if (st != FXC_OK) return st;
    if (stream_capturing(s)) note_captured(p, s);
    const size_t acc_bytes = (size_t)kNumProducts * p->half * sizeof(double);
    const bool phased = (p->desc.products & (FXC_PRODUCT_PHASED | FXC_PRODUCT_RAW_OUT)) != 0;
    if (nf <= 0 || !p->decode_ok[0] || !p->decode_ok[1]) {
        // every frame fails to decode: zero products, zero .raw frames (main.rs:2226-2233)
        HIP_TRY(launch_fill_zero(d_acc, acc_bytes, s));
        if (phased && d_raw_out && nf > 0)
            HIP_TRY(launch_fill_zero(d_raw_out, (size_t)nf * (size_t)plan_bpf_out(p), s));
        return FXC_OK;
    }
    SectorArgs a;
    fill_args(p, a);
    fill_delay(dc, dtab, a.dly);
    a.raw1 = d_raw1;
    a.raw2 = d_raw2;
    a.first_frame = first_frame;
    a.nf = nf;
    if (phased && p->fft_form == 2) return enqueue_phased_large(p, a, nf, d_raw_out, d_acc, s);
    if (phased && phased_fused_ok(p)) return enqueue_phased_fused(p, a, nf, d_raw_out, d_acc, s);
    if (phased) return enqueue_phased(p, a, nf, d_raw_out, d_acc, s);
    if (p->fft_form == 2 || large_fused_ok(p)) return enqueue_large(p, a, nf, d_acc, s);
    const int wg_cap = std::max(1, p->max_wg / p->grid_div);
    // Plain XCF over more than kChainMinOverlap bins: slabs of frames whose
    // f32 phase chains fit the chain scratch; the slabs add into the same
    // workgroup slots in order.
    const bool chain = p->fft_form == 0 && xcf_uses_chain(p);
    const int ld = chain ? chain_ld_for(a.ov) : 0;
    int64_t slab = nf;
    if (chain) {
        slab = std::min<int64_t>(nf, std::max<int64_t>(kFrameBatch, (int64_t)(kChainSlabBytes / ((size_t)ld * 8))));
        GROW(p, p->d_chain, (size_t)slab * ld * 8, s, "phase chains");
    }
    int64_t fpw = (slab + wg_cap - 1) / wg_cap;
    const int batch = p->fft_form == 1 ? 8 : kFrameBatch;   // MFMA form: whole frames per wave
    fpw = ((fpw + batch - 1) / batch) * batch;
    if (fpw < batch) fpw = batch;
    const int n_wg = (int)((slab + fpw - 1) / fpw);
    a.frames_per_wg = fpw;
    a.part = (double *)p->part.p;
    std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
    if (p->timing) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (s && hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone)
            return fail(FXC_ERR_INVALID_ARG, "plan timing is not available inside a graph capture");
        if (!p->pool.empty()) {
            ev = p->pool.back();
            p->pool.pop_back();
        } else {
            HIP_TRY(hipEventCreate(&ev.first));
            HIP_TRY(hipEventCreate(&ev.second));
        }
        HIP_TRY(hipEventRecord(ev.first, s));
    }
    for (int64_t c0 = 0; c0 < nf; c0 += slab) {
        const int64_t cn = std::min(slab, nf - c0);
        SectorArgs b = a;
        b.raw1 = a.raw1 + c0 * p->st[0].bpf;
        b.raw2 = a.raw2 + c0 * p->st[1].bpf;
        b.first_frame = a.first_frame + c0;
        b.nf = cn;
        const int wgs = (int)((cn + fpw - 1) / fpw);
        if (chain) {
            b.chain = (const float2 *)p->d_chain.p;
            b.chain_ld = ld;
            b.accumulate = c0 > 0 ? 1 : 0;
            HIP_TRY(launch_phase_chain(b, CHAIN_XCF, (float2 *)p->d_chain.p, s));
        }
        if (p->fft_form == 1) HIP_TRY(launch_sector_mfma(b, p->desc.precision, wgs, s));
        else HIP_TRY(launch_sector(b, p->desc.precision, wgs, s));
    }
    if (p->timing) {
        HIP_TRY(hipEventRecord(ev.second, s));
        p->pending.push_back(ev);
    }
    HIP_TRY(launch_reduce((const double *)p->part.p, n_wg, p->half, d_acc,
                          (double *)p->red_scratch.p, 64, s));
    return FXC_OK;
}

extern "C" fxc_status fxc_plan_autotune(fxc_plan *p, const uint8_t *d_raw1, const uint8_t *d_raw2,
                                        int64_t nf, const fxc_delay_cfg *delay, void *stream,
                                        fxc_tune_report *rep) {
    if (!p || !d_raw1 || !d_raw2 || !delay || nf <= 0)
        return fail(FXC_ERR_INVALID_ARG, "bad autotune arguments");
    if (p->desc.products & (FXC_PRODUCT_PHASED | FXC_PRODUCT_RAW_OUT))
        return fail(FXC_ERR_UNSUPPORTED, "autotune covers the ACF/XCF sector path");
    hipStream_t s = (hipStream_t)stream;
    const size_t acc_bytes = (size_t)kNumProducts * p->half * sizeof(double);
    if (stream_capturing(s)) return fail(FXC_ERR_INVALID_ARG, "autotune cannot run inside a graph capture");
    if (p->d_acc.ensure(acc_bytes) != hipSuccess) return fail(FXC_ERR_NOMEM, "device allocation failed");
    struct Cand {
        int variant, grid_div, m1;
        int64_t chunk;
    };
    std::vector<Cand> cands;
    // the fused four-step at its default chunk (timed first), twice and half it
    auto add_fused = [p](std::vector<Cand> &cs) {
        const int64_t d = fused_chunk_default(p);
        for (int64_t c : {d, 2 * d, d / 2})
            if (c >= 16 && c <= p->chunk) cs.push_back({1, 1, 128, c});
    };
    if (p->fft_form == 2) {
        // variant 1: the fused four-step (M1 = 128), the default where it
        // applies and so timed first; variant 0: the four-step through HBM
        // at either split
        const int M = p->M;
        if (large_fused_possible(p)) add_fused(cands);
        std::vector<int> m1s = {M / 128};
        if (M / 128 != 128) m1s.push_back(128);
        for (int m1 : m1s)
            for (int64_t c : {p->chunk, p->chunk / 2, p->chunk / 4})
                if (c >= 16) cands.push_back({0, 1, m1, c});
    } else if (p->fft_form == 0) {
        if (large_fused_possible(p)) add_fused(cands);   // N = 16384: the default first
        const int nv = sector_has_variant(p->N) ? 2 : 1;
        for (int v = 0; v < nv; ++v)
            for (int d : {1, 2, 4}) cands.push_back({v, d, 0, 0});
    } else {
        // MFMA form: one workgroup per CU; at N = 1024 the kernel
        // organization varies (1, the default: producer/consumer waves; 0:
        // every wave carries whole frames).  The default is timed first.
        // N = 2048 has one organization.
        if (p->N == 1024)
            for (int v : {1, 2, 0}) cands.push_back({v, 1, 0, 0});
        else
            cands.push_back({1, 1, 0, 0});
    }
    const int save_variant = p->variant, save_div = p->grid_div, save_m1 = p->M1, save_m2 = p->M2;
    const int64_t save_chunk = p->chunk_use;
    const bool save_timing = p->timing;
    p->timing = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    double best = 1e300, dflt = 0.0;
    size_t best_i = 0;
    fxc_status st = FXC_OK;
    for (size_t i = 0; i < cands.size() && st == FXC_OK; ++i) {
        const Cand &c = cands[i];
        p->variant = c.variant;
        p->grid_div = c.grid_div;
        if (c.m1) {
            p->M1 = c.m1;
            p->M2 = p->M / c.m1;
            p->chunk_use = c.chunk;
        }
        double t[3];
        for (int r = 0; r < 4 && st == FXC_OK; ++r) {   // run 0 warms up
            if (hipEventRecord(e0, s) != hipSuccess) st = fail(FXC_ERR_HIP, "event record failed");
            if (st == FXC_OK)
                st = enqueue_sector(p, d_raw1, d_raw2, 0, nf, *delay, (double *)p->d_acc.p, nullptr, s);
            if (st == FXC_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
                st = fail(FXC_ERR_HIP, "event timing failed");
            float ms = 0.f;
            if (st == FXC_OK && r > 0) {
                (void)hipEventElapsedTime(&ms, e0, e1);
                t[r - 1] = ms;
            }
        }
        if (st != FXC_OK) break;
        std::sort(t, t + 3);
        if (i == 0) dflt = t[1];
        if (t[1] < best) {
            best = t[1];
            best_i = i;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    p->timing = save_timing;
    if (st != FXC_OK) {
        p->variant = save_variant;
        p->grid_div = save_div;
        p->M1 = save_m1;
        p->M2 = save_m2;
        p->chunk_use = save_chunk;
        return st;
    }
    const Cand &c = cands[best_i];
    p->variant = c.variant;
    p->grid_div = c.grid_div;
    if (c.m1) {
        p->M1 = c.m1;
        p->M2 = p->M / c.m1;
        p->chunk_use = c.chunk;
    }
    if (rep) {
        const bool four_step = p->fft_form == 2 || large_fused_ok(p);
        rep->candidates = (int32_t)cands.size();
        rep->variant = p->variant;
        rep->grid_divisor = p->grid_div;
        rep->split_m1 = four_step ? p->M1 : 0;
        rep->chunk_frames = four_step ? p->chunk_use : 0;
        rep->best_ms = best;
        rep->default_ms = dflt;
    }
    return FXC_OK;
}

extern "C" fxc_status fxc_plan_set_tuning(fxc_plan *p, const fxc_tune_report *c) {
    if (!p || !c) return fail(FXC_ERR_INVALID_ARG, "null argument");
    const bool v1_ok = (p->fft_form == 0 && sector_has_variant(p->N)) || p->fft_form == 1 ||
                       ((large_fused_possible(p) || phased_fused_possible(p)) && c->split_m1 == 128);
    const bool v0_ok = !(p->fft_form == 1 && p->N != 1024);
    const bool v2_ok = p->fft_form == 1 && p->N == 1024;
    if ((c->variant == 1 && !v1_ok) || (c->variant == 0 && !v0_ok) || (c->variant == 2 && !v2_ok) ||
        c->variant < 0 || c->variant > 2)
        return fail(FXC_ERR_INVALID_ARG, "variant %d not available for this plan", c->variant);
    if (c->grid_divisor != 1 && c->grid_divisor != 2 && c->grid_divisor != 4)
        return fail(FXC_ERR_INVALID_ARG, "grid_divisor must be 1, 2 or 4");
    if (p->fft_form == 2) {
        if ((c->split_m1 != 128 && c->split_m1 != 256) || p->M % c->split_m1 != 0 ||
            (p->M / c->split_m1 != 128 && p->M / c->split_m1 != 256))
            return fail(FXC_ERR_INVALID_ARG, "split_m1 %d invalid for M = %d", c->split_m1, p->M);
        if (c->chunk_frames < 1 || c->chunk_frames > p->chunk)
            return fail(FXC_ERR_INVALID_ARG, "chunk_frames outside 1..%lld", (long long)p->chunk);
        p->M1 = c->split_m1;
        p->M2 = p->M / c->split_m1;
        p->chunk_use = c->chunk_frames;
    } else if (c->variant == 1 && (large_fused_possible(p) || phased_fused_possible(p))) {
        // N = 16384: the fused four-step
        if (c->chunk_frames < 1 || c->chunk_frames > p->chunk)
            return fail(FXC_ERR_INVALID_ARG, "chunk_frames outside 1..%lld", (long long)p->chunk);
        p->chunk_use = c->chunk_frames;
    }
    p->variant = c->variant;
    p->grid_div = c->grid_divisor;
    return FXC_OK;
}

extern "C" fxc_status fxc_sector_run_device(fxc_plan *p, const uint8_t *d_raw1,
                                            const uint8_t *d_raw2, int64_t first_frame, int64_t nf,
                                            const fxc_delay_cfg *delay, double *d_acc,
                                            uint8_t *d_raw_out, void *stream) {
    if (!p || !delay || !d_acc) return fail(FXC_ERR_INVALID_ARG, "null argument");
    if (nf < 0) return fail(FXC_ERR_INVALID_ARG, "nf < 0");
    if (nf > 0 && (!d_raw1 || !d_raw2)) return fail(FXC_ERR_INVALID_ARG, "null raw buffer");
    hipStream_t s = (hipStream_t)stream;  // NULL: the HIP null stream
    return enqueue_sector(p, d_raw1, d_raw2, first_frame, nf, *delay, d_acc, d_raw_out, s);
}

// The sector path of fxc_sector_run_device with every scratch buffer in
// caller-given device memory (fxc_noalloc.hip passes its static array):
// [geometric table | reduction chunk sums | workgroup partial slots].  Only
// the paths whose whole scratch is the partial slots run here: ACF/XCF plans
// of the fused kernels without the plain form's phase-chain table.  The
// grid is the plan's, capped by the slots that fit.  Allocates nothing and
// never synchronises.
extern "C" fxc_status fxc_internal_sector_noalloc(fxc_plan *p, const uint8_t *d_raw1,
                                                  const uint8_t *d_raw2, int64_t first_frame,
                                                  int64_t nf, const fxc_delay_cfg *delay,
                                                  double *d_acc, void *stream, uint8_t *scratch,
                                                  size_t scratch_bytes, size_t table_bytes,
                                                  int device) {
    if (!p || !delay || !d_acc || !scratch) return fail(FXC_ERR_INVALID_ARG, "null argument");
    if (nf < 0) return fail(FXC_ERR_INVALID_ARG, "nf < 0");
    if (nf > 0 && (!d_raw1 || !d_raw2)) return fail(FXC_ERR_INVALID_ARG, "null raw buffer");
    if (device != p->device)
        return fail(FXC_ERR_INVALID_ARG, "the plan was created on device %d but device %d is current",
                    p->device, device);
    if (p->desc.products & (FXC_PRODUCT_PHASED | FXC_PRODUCT_RAW_OUT))
        return fail(FXC_ERR_UNSUPPORTED,
                    "the allocation-free entry runs ACF/XCF plans; phased plans keep per-chunk "
                    "spectra and rotation chains in plan scratch (use fxc_sector_run_device)");
    if (p->fft_form == 2)
        return fail(FXC_ERR_UNSUPPORTED,
                    "fft_len %d runs its four-step FFT through per-plan chunk scratch "
                    "(use fxc_sector_run_device)", p->N);
    if (xcf_uses_chain(p))
        return fail(FXC_ERR_UNSUPPORTED,
                    "the plain form's XCF over %lld overlap bins replays the reference's f32 phase "
                    "chains through a per-plan table (use fxc_sector_run_device, or the chosen form)",
                    (long long)(p->desc.a1e - p->desc.a1s));
    hipStream_t s = (hipStream_t)stream;
    const double *dtab = nullptr;
    if (delay->geom_table && delay->geom_table_len > 0) {
        const size_t tb = (size_t)delay->geom_table_len * sizeof(double);
        if (tb > table_bytes)
            return fail(FXC_ERR_INVALID_ARG,
                        "geometric table of %lld entries exceeds the allocation-free entry's limit "
                        "of %zu entries", (long long)delay->geom_table_len, table_bytes / sizeof(double));
        if (stream_capturing(s))
            return fail(FXC_ERR_INVALID_ARG,
                        "the allocation-free entry copies the geometric table from host memory on "
                        "each call, which a graph capture cannot hold; capture fxc_sector_run_device");
        HIP_TRY(hipMemcpyAsync(scratch, delay->geom_table, tb, hipMemcpyHostToDevice, s));
        dtab = (const double *)scratch;
    }
    const size_t acc_bytes = (size_t)kNumProducts * p->half * sizeof(double);
    if (nf <= 0 || !p->decode_ok[0] || !p->decode_ok[1]) {
        HIP_TRY(launch_fill_zero(d_acc, acc_bytes, s));
        return FXC_OK;
    }
    const int kReduceChunks = 64;
    const size_t slot = acc_bytes;
    // per workgroup: its partial slot and, for the fused N = 1024 / 2048
    // form, its consumers' FP64 totals
    const size_t per_wg = slot + (p->fft_form == 1 ? kMfConsumerAccBytes : 0);
    const size_t head = table_bytes + (size_t)kReduceChunks * slot + 256;   // + alignment of the totals
    const int64_t slots = scratch_bytes > head ? (int64_t)((scratch_bytes - head) / per_wg) : 0;
    if (slots < 1)
        return fail(FXC_ERR_UNSUPPORTED, "fft_len %d: %zu-byte partial slots do not fit the %zu-byte scratch",
                    p->N, slot, scratch_bytes);
    double *red = (double *)(scratch + table_bytes);
    double *part = red + (size_t)kReduceChunks * kNumProducts * p->half;
    const uintptr_t cacc_at = ((uintptr_t)(part + (size_t)slots * kNumProducts * p->half) + 255) & ~(uintptr_t)255;
    double *cacc = (double *)cacc_at;
    SectorArgs a;
    fill_args(p, a);
    a.cons_acc = p->fft_form == 1 ? cacc : nullptr;
    fill_delay(*delay, dtab, a.dly);
    a.raw1 = d_raw1;
    a.raw2 = d_raw2;
    a.first_frame = first_frame;
    a.nf = nf;
    const int64_t wg_cap = std::min<int64_t>(std::max(1, p->max_wg / p->grid_div), slots);
    int64_t fpw = (nf + wg_cap - 1) / wg_cap;
    const int batch = p->fft_form == 1 ? 8 : kFrameBatch;
    fpw = ((fpw + batch - 1) / batch) * batch;
    if (fpw < batch) fpw = batch;
    const int n_wg = (int)((nf + fpw - 1) / fpw);
    a.frames_per_wg = fpw;
    a.part = part;
    if (p->fft_form == 1) HIP_TRY(launch_sector_mfma(a, p->desc.precision, n_wg, s));
    else HIP_TRY(launch_sector(a, p->desc.precision, n_wg, s));
    HIP_TRY(launch_reduce(part, n_wg, p->half, d_acc, red, kReduceChunks, s));
    return FXC_OK;
}

// One chunk of the whole-scan cross spectrum (main.rs:1760-1849: the XCF
// of nf frames on the plan's output grid), scaled by 1/N^2 and added into
// ac (main.rs:1751, 1850).  Frames that fail to decode add nothing.
static fxc_status fringe_check(const fxc_plan *p) {
    if (!(p->desc.products & FXC_PRODUCT_XCF) || (p->desc.products & (FXC_PRODUCT_PHASED | FXC_PRODUCT_RAW_OUT)))
        return fail(FXC_ERR_UNSUPPORTED, "the whole-scan cross spectrum needs an XCF plan without phased products");
    return FXC_OK;
}

extern "C" fxc_status fxc_fringe_accumulate_device(fxc_plan *p, const uint8_t *d_raw1,
                                                   const uint8_t *d_raw2, int64_t first_frame,
                                                   int64_t nf, const fxc_delay_cfg *delay,
                                                   double *d_ac, void *stream) {
    if (!p || !delay || !d_ac) return fail(FXC_ERR_INVALID_ARG, "null argument");
    fxc_status st = fringe_check(p);
    if (st != FXC_OK) return st;
    if (nf < 0) return fail(FXC_ERR_INVALID_ARG, "nf < 0");
    if (nf > 0 && (!d_raw1 || !d_raw2)) return fail(FXC_ERR_INVALID_ARG, "null raw buffer");
    const size_t acc_bytes = (size_t)kNumProducts * p->half * sizeof(double);
    hipStream_t s = (hipStream_t)stream;
    GROW(p, p->d_fringe, acc_bytes, s, "whole-scan spectra");
    double *d = (double *)p->d_fringe.p;
    st = enqueue_sector(p, d_raw1, d_raw2, first_frame, nf, *delay, d, nullptr, s);
    if (st != FXC_OK) return st;
    const double inv_fft2 = 1.0 / ((double)p->N * (double)p->N);
    HIP_TRY(launch_scale_add_xcf(d_ac, d + (size_t)P_12R * p->half, d + (size_t)P_12I * p->half, p->half,
                                 inv_fft2, s));
    return FXC_OK;
}

extern "C" fxc_status fxc_sector_run(fxc_plan *p, const uint8_t *raw1, size_t len1,
                                     const uint8_t *raw2, size_t len2, int64_t first_frame,
                                     int64_t nf, const fxc_delay_cfg *delay, double *acc_ph,
                                     double *acc11, double *acc12, double *acc22,
                                     uint8_t *raw_out, int64_t *n_failed) {
    if (!p || !delay) return fail(FXC_ERR_INVALID_ARG, "null argument");
    if (nf < 0) return fail(FXC_ERR_INVALID_ARG, "nf < 0");
    const size_t need1 = (size_t)nf * p->st[0].bpf, need2 = (size_t)nf * p->st[1].bpf;
    if (nf > 0 && (!raw1 || !raw2 || len1 < need1 || len2 < need2))
        return fail(FXC_ERR_INVALID_ARG, "raw buffers shorter than nf*bpf (%zu/%zu < %zu/%zu)", len1,
                    len2, need1, need2);
    hipStream_t s = p->stream;
    const size_t acc_bytes = (size_t)kNumProducts * p->half * sizeof(double);
    if (p->d_acc.ensure(acc_bytes) != hipSuccess) return fail(FXC_ERR_NOMEM, "device allocation failed");
    const bool want_raw = raw_out && (p->desc.products & FXC_PRODUCT_RAW_OUT) && nf > 0;
    const size_t raw_bytes = want_raw ? (size_t)nf * (size_t)plan_bpf_out(p) : 0;
    if (want_raw && p->d_raw_out.ensure(raw_bytes) != hipSuccess)
        return fail(FXC_ERR_NOMEM, "device allocation of %zu bytes failed", raw_bytes);
    if (nf > 0) {
        if (p->d_raw[0].ensure(need1) != hipSuccess || p->d_raw[1].ensure(need2) != hipSuccess)
            return fail(FXC_ERR_NOMEM, "device allocation of %zu bytes failed", need1 + need2);
        HIP_TRY(hipMemcpyAsync(p->d_raw[0].p, raw1, need1, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(p->d_raw[1].p, raw2, need2, hipMemcpyHostToDevice, s));
    }
    fxc_status st = enqueue_sector(p, (const uint8_t *)p->d_raw[0].p, (const uint8_t *)p->d_raw[1].p,
                                   first_frame, nf, *delay, (double *)p->d_acc.p,
                                   want_raw ? (uint8_t *)p->d_raw_out.p : nullptr, s);
    if (st != FXC_OK) return st;
    std::vector<double> host((size_t)kNumProducts * p->half);
    HIP_TRY(hipMemcpyAsync(host.data(), p->d_acc.p, acc_bytes, hipMemcpyDeviceToHost, s));
    if (want_raw) HIP_TRY(hipMemcpyAsync(raw_out, p->d_raw_out.p, raw_bytes, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (p->timing) {
        fxc_status ts = collect_timing(p);
        if (ts != FXC_OK) return ts;
    }
    st = read_error_word(p, s);
    if (st != FXC_OK) return st;
    const int h = p->half;
    if (acc_ph) memcpy(acc_ph, host.data() + P_PH * h, h * sizeof(double));
    if (acc11) memcpy(acc11, host.data() + P_11 * h, h * sizeof(double));
    if (acc22) memcpy(acc22, host.data() + P_22 * h, h * sizeof(double));
    if (acc12)
        for (int k = 0; k < h; ++k) {
            acc12[2 * k] = host[P_12R * h + k];
            acc12[2 * k + 1] = host[P_12I * h + k];
        }
    if (n_failed) *n_failed = (p->decode_ok[0] && p->decode_ok[1]) ? 0 : nf;
    return FXC_OK;
}

extern "C" fxc_status fxc_fringe_accumulate(fxc_plan *p, const uint8_t *raw1, size_t len1,
                                            const uint8_t *raw2, size_t len2, int64_t first_frame,
                                            int64_t nf, const fxc_delay_cfg *delay, double *ac,
                                            int64_t *n_failed) {
    if (!p || !delay || !ac) return fail(FXC_ERR_INVALID_ARG, "null argument");
    fxc_status st = fringe_check(p);
    if (st != FXC_OK) return st;
    const size_t h = (size_t)p->half;
    std::vector<double> ph(h), a11(h), a22(h), a12(2 * h);
    st = fxc_sector_run(p, raw1, len1, raw2, len2, first_frame, nf, delay, ph.data(), a11.data(),
                        a12.data(), a22.data(), nullptr, n_failed);
    if (st != FXC_OK) return st;
    const double inv_fft2 = 1.0 / ((double)p->N * (double)p->N);
    for (size_t k = 0; k < h; ++k) {
        ac[2 * k] += a12[2 * k] * inv_fft2;
        ac[2 * k + 1] += a12[2 * k + 1] * inv_fft2;
    }
    return FXC_OK;
}

extern "C" fxc_status fxc_plan_set_timing(fxc_plan *p, int32_t enable) {
    if (!p) return fail(FXC_ERR_INVALID_ARG, "null plan");
    p->timing = enable != 0;
    p->timed_ms = 0.0;
    p->timed_launches = 0;
    return FXC_OK;
}

extern "C" fxc_status fxc_plan_get_timing(fxc_plan *p, double *total_ms, int64_t *launches) {
    if (!p || !total_ms || !launches) return fail(FXC_ERR_INVALID_ARG, "null argument");
    fxc_status st = collect_timing(p);
    if (st != FXC_OK) return st;
    *total_ms = p->timed_ms;
    *launches = p->timed_launches;
    return FXC_OK;
}

// ================================================== per-function entries
extern "C" fxc_status fxc_decode_frames(fxc_plan *p, int32_t station, const uint8_t *raw,
                                        size_t len, int64_t nframes, const int64_t *shifts,
                                        float *out, int64_t *n_failed) {
    if (!p || !out || (station != 0 && station != 1) || nframes < 0)
        return fail(FXC_ERR_INVALID_ARG, "bad decode arguments");
    const StationDev &sd = p->st[station];
    const size_t need = (size_t)nframes * sd.bpf;
    if (nframes > 0 && (!raw || len < need)) return fail(FXC_ERR_INVALID_ARG, "raw shorter than nframes*bpf");
    if (n_failed) *n_failed = sd.decode_ok ? 0 : nframes;
    const size_t out_bytes = (size_t)nframes * p->N * sizeof(float);
    if (nframes == 0) return FXC_OK;
    if (!sd.decode_ok) {
        memset(out, 0, out_bytes);
        return FXC_OK;
    }
    hipStream_t s = p->stream;
    if (p->d_tmp_in.ensure(need) != hipSuccess || p->d_tmp_out.ensure(out_bytes) != hipSuccess)
        return fail(FXC_ERR_NOMEM, "device allocation failed");
    HIP_TRY(hipMemcpyAsync(p->d_tmp_in.p, raw, need, hipMemcpyHostToDevice, s));
    const int64_t *dsh = nullptr;
    if (shifts) {
        if (p->d_shifts.ensure((size_t)nframes * 8) != hipSuccess)
            return fail(FXC_ERR_NOMEM, "device allocation failed");
        HIP_TRY(hipMemcpyAsync(p->d_shifts.p, shifts, (size_t)nframes * 8, hipMemcpyHostToDevice, s));
        dsh = (const int64_t *)p->d_shifts.p;
    }
    SectorArgs a;
    fill_args(p, a);
    a.nf = nframes;
    if (station == 0) a.raw1 = (const uint8_t *)p->d_tmp_in.p;
    else a.raw2 = (const uint8_t *)p->d_tmp_in.p;
    HIP_TRY(launch_decode_frames(a, station, dsh, 1, p->d_tmp_out.p, false, s));
    HIP_TRY(hipMemcpyAsync(out, p->d_tmp_out.p, out_bytes, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return FXC_OK;
}

static fxc_status run_fft(fxc_plan *p, const float *in, int64_t nframes, float *out, bool inverse) {
    if (!p || !in || !out || nframes < 0) return fail(FXC_ERR_INVALID_ARG, "bad fft arguments");
    if (nframes == 0) return FXC_OK;
    const size_t real_bytes = (size_t)nframes * p->N * sizeof(float);
    const size_t cplx_bytes = (size_t)nframes * p->half * 2 * sizeof(float);
    const size_t in_bytes = inverse ? cplx_bytes : real_bytes;
    const size_t out_bytes = inverse ? real_bytes : cplx_bytes;
    hipStream_t s = p->stream;
    if (p->d_tmp_in.ensure(in_bytes) != hipSuccess || p->d_tmp_out.ensure(out_bytes) != hipSuccess)
        return fail(FXC_ERR_NOMEM, "device allocation failed");
    HIP_TRY(hipMemcpyAsync(p->d_tmp_in.p, in, in_bytes, hipMemcpyHostToDevice, s));
    SectorArgs a;
    fill_args(p, a);
    if (p->fft_form == 2 && inverse) {
        for (int64_t c0 = 0; c0 < nframes; c0 += p->chunk) {   // through the chunk scratch
            const int64_t cn = std::min(p->chunk, nframes - c0);
            HIP_TRY(launch_irfft_large((const float2 *)p->d_tmp_in.p + c0 * p->half, cn, p->M1, p->M2,
                                       (const float2 *)(p->M1 == 128 ? p->tw_l1.p : p->tw_l2.p),
                                       (const float2 *)(p->M2 == 128 ? p->tw_l1.p : p->tw_l2.p),
                                       (const float2 *)p->tw_lm.p, (const float2 *)p->tw_post.p,
                                       (float2 *)p->d_x[0].p, (float2 *)p->d_y.p, (float2 *)p->d_z[0].p,
                                       false, (float *)p->d_tmp_out.p + c0 * p->N, s));
        }
    } else if (p->fft_form == 2) {
        for (int64_t c0 = 0; c0 < nframes; c0 += p->chunk) {   // through the chunk scratch
            const int64_t cn = std::min(p->chunk, nframes - c0);
            HIP_TRY(launch_fft_large((const float *)p->d_tmp_in.p + c0 * p->N, false, (float2 *)p->d_y.p,
                                     (float2 *)p->d_z[0].p, cn, p->M1, p->M2,
                                     (const float2 *)(p->M1 == 128 ? p->tw_l1.p : p->tw_l2.p),
                                     (const float2 *)(p->M2 == 128 ? p->tw_l1.p : p->tw_l2.p),
                                     (const float2 *)p->tw_lm.p,
                                     p->mf_b1 ? p->mf_tab.p : nullptr, s));
            HIP_TRY(launch_untangle_large((const float2 *)p->d_z[0].p, cn, p->M,
                                          (const float2 *)p->tw_post.p,
                                          (float2 *)p->d_tmp_out.p + c0 * p->half, s));
        }
    } else if (inverse)
        HIP_TRY(launch_irfft_frames(a, (const float2 *)p->d_tmp_in.p, nframes, (float *)p->d_tmp_out.p,
                                    false, s));
    else if (p->fft_form == 1 && p->N == 1024)
        HIP_TRY(launch_rfft_mfma(a, (const float *)p->d_tmp_in.p, nframes, (float2 *)p->d_tmp_out.p, s));
    else {
        if (p->fft_form == 1) a.mf_tab = nullptr;   // N = 2048: its table is the fused kernel's
        HIP_TRY(launch_rfft_frames(a, (const float *)p->d_tmp_in.p, nframes, (float2 *)p->d_tmp_out.p, s));
    }
    HIP_TRY(hipMemcpyAsync(out, p->d_tmp_out.p, out_bytes, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return FXC_OK;
}

extern "C" fxc_status fxc_rfft_frames(fxc_plan *p, const float *in, int64_t nframes, float *out) {
    return run_fft(p, in, nframes, out, false);
}
extern "C" fxc_status fxc_irfft_frames(fxc_plan *p, const float *in, int64_t nframes, float *out) {
    return run_fft(p, in, nframes, out, true);
}

extern "C" fxc_status fxc_delay_eval(const fxc_delay_cfg *cfg, int64_t first, int64_t n,
                                     fxc_frame_delay *out) {
    static_assert(sizeof(fxc_frame_delay) == sizeof(FrameDelayDev), "layout mismatch");
    if (!cfg || !out || n < 0) return fail(FXC_ERR_INVALID_ARG, "bad delay_eval arguments");
    if (n == 0) return FXC_OK;
    int dc = 0;
    if (hipGetDeviceCount(&dc) != hipSuccess || dc < 1) return fail(FXC_ERR_HIP, "no HIP device");
    void *dtab = nullptr, *dout = nullptr;
    hipStream_t s = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    fxc_status rc = FXC_OK;
    do {
        if (cfg->geom_table && cfg->geom_table_len > 0) {
            if (hipMalloc(&dtab, (size_t)cfg->geom_table_len * 8) != hipSuccess) {
                rc = fail(FXC_ERR_NOMEM, "table allocation failed");
                break;
            }
            if (hipMemcpy(dtab, cfg->geom_table, (size_t)cfg->geom_table_len * 8,
                          hipMemcpyHostToDevice) != hipSuccess) {
                rc = fail(FXC_ERR_HIP, "table upload failed");
                break;
            }
        }
        if (hipMalloc(&dout, (size_t)n * sizeof(FrameDelayDev)) != hipSuccess) {
            rc = fail(FXC_ERR_NOMEM, "output allocation failed");
            break;
        }
        DelayDev d;
        fill_delay(*cfg, (const double *)dtab, d);
        hipError_t e = launch_delay_eval(d, first, n, (FrameDelayDev *)dout, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(out, dout, (size_t)n * sizeof(FrameDelayDev), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = fail(FXC_ERR_HIP, "delay_eval: %s", hipGetErrorString(e));
    } while (0);
    if (dtab) (void)hipFree(dtab);
    if (dout) (void)hipFree(dout);
    (void)hipStreamDestroy(s);
    return rc;
}

extern "C" fxc_status fxc_quantise_frames(fxc_plan *p, const float *in, int64_t nframes,
                                          uint8_t *out) {
    if (!p || !in || !out || nframes < 0) return fail(FXC_ERR_INVALID_ARG, "bad quantise arguments");
    if (p->desc.bit_out < 1 || p->desc.bit_out > 16 || ((int64_t)p->N * p->desc.bit_out) % 32 != 0)
        return fail(FXC_ERR_UNSUPPORTED, "fft_len*bit_out must fill whole 32-bit words (bit_out 1..16)");
    if (p->desc.st[0].nlevels > 256) return fail(FXC_ERR_UNSUPPORTED, "at most 256 output levels");
    if (nframes == 0) return FXC_OK;
    const size_t bpo = (size_t)p->N * p->desc.bit_out / 8;
    const size_t in_bytes = (size_t)nframes * p->N * sizeof(float), out_bytes = (size_t)nframes * bpo;
    hipStream_t s = p->stream;
    if (p->d_tmp_in.ensure(in_bytes) != hipSuccess || p->d_tmp_out.ensure(out_bytes) != hipSuccess)
        return fail(FXC_ERR_NOMEM, "device allocation failed");
    HIP_TRY(hipMemcpyAsync(p->d_tmp_in.p, in, in_bytes, hipMemcpyHostToDevice, s));
    SectorArgs a;
    fill_args(p, a);
    HIP_TRY(launch_quantise_frames(a, (const float *)p->d_tmp_in.p, nframes, (uint8_t *)p->d_tmp_out.p, s));
    HIP_TRY(hipMemcpyAsync(out, p->d_tmp_out.p, out_bytes, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return FXC_OK;
}

// ============================================================ synthetic input
extern "C" fxc_status fxc_synth_station_device(uint8_t *d_out, int64_t nsamples, int32_t bits,
                                               const int32_t shuffle_internal[32], int32_t lsb,
                                               uint64_t seed, int32_t station, double rho,
                                               int64_t delay_samples, void *stream) {
    if (!d_out || !shuffle_internal || nsamples < 0 || (station != 0 && station != 1))
        return fail(FXC_ERR_INVALID_ARG, "bad synth arguments");
    if (bits < 1 || bits > 16 || 32 % bits != 0 || (nsamples * bits) % 32 != 0)
        return fail(FXC_ERR_UNSUPPORTED, "synth needs bits dividing 32 and whole words");
    if (!(rho >= 0.0 && rho <= 1.0)) return fail(FXC_ERR_INVALID_ARG, "rho outside [0, 1]");
    int seen[32] = {0};
    for (int i = 0; i < 32; ++i)
        if (shuffle_internal[i] < 0 || shuffle_internal[i] > 31 || seen[shuffle_internal[i]]++)
            return fail(FXC_ERR_INVALID_ARG, "shuffle is not a permutation of 0..31");
    if (nsamples == 0) return FXC_OK;
    HIP_TRY(launch_synth(d_out, nsamples, bits, shuffle_internal, lsb, seed, station, rho,
                         delay_samples, (hipStream_t)stream));
    return FXC_OK;
}

extern "C" fxc_status fxc_synth_spectral_pair_device(const fxc_synth_spectral_desc *d,
                                                     int64_t first_frame, int64_t nframes,
                                                     uint8_t *d_raw1, uint8_t *d_raw2,
                                                     void *stream) {
    if (!d || !d_raw1 || !d_raw2 || nframes < 0 || first_frame < 0)
        return fail(FXC_ERR_INVALID_ARG, "bad spectral synth arguments");
    const int N = d->fft_len;
    if (!is_pow2(N) || N < 64 || N > 65536)
        return fail(FXC_ERR_UNSUPPORTED, "spectral synth: fft_len %d is not a power of two 64..65536", N);
    for (int i = 0; i < 2; ++i) {
        const int b = d->bits[i];
        if (b != 1 && b != 2 && b != 4 && b != 8)
            return fail(FXC_ERR_UNSUPPORTED, "spectral synth: bits must be 1, 2, 4 or 8");
        int seen[32] = {0};
        for (int k = 0; k < 32; ++k)
            if (d->shuffle_internal[i][k] < 0 || d->shuffle_internal[i][k] > 31 ||
                seen[d->shuffle_internal[i][k]]++)
                return fail(FXC_ERR_INVALID_ARG, "shuffle is not a permutation of 0..31");
    }
    if (!(d->rho_continuum >= 0.0 && d->rho_continuum < 1.0) || !(d->rho_line < 1.0) ||
        std::llabs(d->shift_bins) > N / 2 || (d->rho_line > 0.0 && !(d->line_fwhm_bins > 0.0)))
        return fail(FXC_ERR_INVALID_ARG, "spectral synth: coherence outside [0, 1), shift beyond "
                                         "the half spectrum or line without width");
    if ((first_frame + nframes) > ((int64_t)1 << 38))
        return fail(FXC_ERR_INVALID_ARG, "spectral synth: frame index beyond 2^38");
    if (nframes == 0) return FXC_OK;
    hipStream_t s = (hipStream_t)stream;
    // The inverse FFT runs through a private plan of this length (its
    // twiddles and, at N >= 32768, its chunk scratch).
    fxc_plan_desc pd;
    memset(&pd, 0, sizeof pd);
    pd.fft_len = N;
    static const double lv1[2] = {-1.0, 1.0};
    for (int i = 0; i < 2; ++i) {
        pd.st[i].bits = 1;
        for (int k = 0; k < 32; ++k) pd.st[i].shuffle_internal[k] = k;
        pd.st[i].levels = lv1;
        pd.st[i].nlevels = 2;
    }
    pd.a1s = 0;
    pd.a1e = N / 2;
    pd.a2s = 0;
    pd.df_hz = 1.0;
    pd.products = FXC_PRODUCT_ACF;
    pd.w1 = pd.w2 = 1.0;
    pd.bit_out = 1;
    pd.precision = FXC_PRECISION_PLAIN_FP32;
    fxc_plan *p = nullptr;
    fxc_status st = fxc_plan_create(&pd, &p);
    if (st != FXC_OK) return st;
    struct Guard {
        fxc_plan *p;
        hipStream_t s;
        ~Guard() {
            (void)hipStreamSynchronize(s);
            fxc_plan_destroy(p);
        }
    } guard{p, s};
    const int64_t chunk = p->fft_form == 2
                              ? p->chunk
                              : std::max<int64_t>(1, ((int64_t)256 << 20) / ((int64_t)N * 4));
    if (p->d_tmp_in.ensure((size_t)chunk * p->half * 8) != hipSuccess ||
        p->d_tmp_out.ensure((size_t)chunk * N * 4) != hipSuccess)
        return fail(FXC_ERR_NOMEM, "spectral synth: FFT scratch allocation failed");
    SectorArgs a;
    fill_args(p, a);
    uint8_t *outs[2] = {d_raw1, d_raw2};
    for (int stn = 0; stn < 2; ++stn) {
        const int64_t bpf = (int64_t)N * d->bits[stn] / 8;
        for (int64_t c0 = 0; c0 < nframes; c0 += chunk) {
            const int64_t cn = std::min(chunk, nframes - c0);
            float2 *spec = (float2 *)p->d_tmp_in.p;
            float *x = (float *)p->d_tmp_out.p;
            HIP_TRY(launch_synth_spectrum(spec, first_frame + c0, cn, p->M, stn, d->shift_bins, d->seed,
                                          d->rho_continuum, d->rho_line > 0.0 ? d->rho_line : 0.0,
                                          d->line_bin, d->line_fwhm_bins, s));
            if (p->fft_form == 2)
                HIP_TRY(launch_irfft_large(spec, cn, p->M1, p->M2,
                                           (const float2 *)(p->M1 == 128 ? p->tw_l1.p : p->tw_l2.p),
                                           (const float2 *)(p->M2 == 128 ? p->tw_l1.p : p->tw_l2.p),
                                           (const float2 *)p->tw_lm.p, (const float2 *)p->tw_post.p,
                                           (float2 *)p->d_x[0].p, (float2 *)p->d_y.p,
                                           (float2 *)p->d_z[0].p, false, x, s));
            else
                HIP_TRY(launch_irfft_frames(a, spec, cn, x, false, s));
            HIP_TRY(launch_synth_pack(x, cn * N, d->bits[stn], d->lsb[stn], d->shuffle_internal[stn],
                                      d->thresh_sigma, d->clip_sigma, outs[stn] + c0 * bpf, s));
        }
    }
    HIP_TRY(hipStreamSynchronize(s));
    return FXC_OK;
}

// ============================================================ host pipeline
// Streaming host input (the reader-thread + per-sector loop of
// main.rs:2098-2213): sectors submitted from host memory are copied on a
// copy stream into one of `depth` device slots while earlier sectors run on
// the compute stream, and each sector's spectra come back into the caller's
// memory.  With page-locked input (fxc_host_alloc) the copies run at PCIe
// rate and overlap the kernels.
extern "C" fxc_status fxc_host_alloc(size_t bytes, void **out) {
    if (!out) return fail(FXC_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    HIP_TRY(hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault));
    return FXC_OK;
}

extern "C" fxc_status fxc_host_free(void *p) {
    if (p) HIP_TRY(hipHostFree(p));
    return FXC_OK;
}

struct fxc_pipeline {
    fxc_plan *plan = nullptr;
    int64_t max_frames = 0;
    int depth = 0;
    hipStream_t copy = nullptr, compute = nullptr;
    struct Slot {
        DevBuf raw[2], acc;
        hipEvent_t copied = nullptr, done = nullptr;
        int32_t *err = nullptr;   // page-locked copy of the plan's error word after this sector
        int64_t ticket = -1;
    };
    std::vector<Slot> slots;
    int64_t next_ticket = 0;
};

static void pipeline_release(fxc_pipeline *pl) {
    if (!pl) return;
    if (pl->copy) (void)hipStreamSynchronize(pl->copy);
    if (pl->compute) (void)hipStreamSynchronize(pl->compute);
    for (auto &s : pl->slots) {
        s.raw[0].release();
        s.raw[1].release();
        s.acc.release();
        if (s.copied) (void)hipEventDestroy(s.copied);
        if (s.done) (void)hipEventDestroy(s.done);
        if (s.err) (void)hipHostFree(s.err);
    }
    if (pl->copy) (void)hipStreamDestroy(pl->copy);
    if (pl->compute) (void)hipStreamDestroy(pl->compute);
    delete pl;
}

extern "C" fxc_status fxc_pipeline_create(fxc_plan *p, int64_t max_frames, int32_t depth,
                                          fxc_pipeline **out) {
    if (!p || !out || max_frames < 1 || depth < 1 || depth > 8)
        return fail(FXC_ERR_INVALID_ARG, "pipeline needs a plan, max_frames >= 1 and depth 1..8");
    if (p->desc.products & FXC_PRODUCT_RAW_OUT)
        return fail(FXC_ERR_UNSUPPORTED, "the pipeline returns spectra; .raw output uses fxc_sector_run");
    *out = nullptr;
    auto *pl = new fxc_pipeline();
    pl->plan = p;
    pl->max_frames = max_frames;
    pl->depth = depth;
    pl->slots.resize((size_t)depth);
    auto bail = [&](fxc_status st) { pipeline_release(pl); return st; };
    if (hipStreamCreateWithFlags(&pl->copy, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&pl->compute, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(FXC_ERR_HIP, "stream creation failed"));
    const size_t acc_bytes = (size_t)kNumProducts * p->half * sizeof(double);
    for (auto &s : pl->slots) {
        if (s.raw[0].ensure((size_t)max_frames * p->st[0].bpf) != hipSuccess ||
            s.raw[1].ensure((size_t)max_frames * p->st[1].bpf) != hipSuccess ||
            s.acc.ensure(acc_bytes) != hipSuccess)
            return bail(fail(FXC_ERR_NOMEM, "pipeline slot allocation failed"));
        if (hipEventCreateWithFlags(&s.copied, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc((void **)&s.err, sizeof(int32_t), hipHostMallocDefault) != hipSuccess)
            return bail(fail(FXC_ERR_HIP, "event or host allocation failed"));
        *s.err = 0;
    }
    *out = pl;
    return FXC_OK;
}

extern "C" fxc_status fxc_pipeline_submit(fxc_pipeline *pl, const uint8_t *raw1, const uint8_t *raw2,
                                          int64_t first_frame, int64_t nf, const fxc_delay_cfg *delay,
                                          double *acc_out, int64_t *ticket) {
    if (!pl || !delay || !acc_out || !ticket) return fail(FXC_ERR_INVALID_ARG, "null argument");
    if (nf < 0 || nf > pl->max_frames)
        return fail(FXC_ERR_INVALID_ARG, "nf %lld outside 0..%lld", (long long)nf, (long long)pl->max_frames);
    if (nf > 0 && (!raw1 || !raw2)) return fail(FXC_ERR_INVALID_ARG, "null raw buffer");
    fxc_plan *p = pl->plan;
    const int64_t t = pl->next_ticket;
    auto &s = pl->slots[(size_t)(t % pl->depth)];
    // the slot's previous sector must have finished reading its input
    if (s.ticket >= 0) HIP_TRY(hipStreamWaitEvent(pl->copy, s.done, 0));
    const size_t n1 = (size_t)nf * p->st[0].bpf, n2 = (size_t)nf * p->st[1].bpf;
    if (nf > 0) {
        HIP_TRY(hipMemcpyAsync(s.raw[0].p, raw1, n1, hipMemcpyHostToDevice, pl->copy));
        HIP_TRY(hipMemcpyAsync(s.raw[1].p, raw2, n2, hipMemcpyHostToDevice, pl->copy));
    }
    HIP_TRY(hipEventRecord(s.copied, pl->copy));
    HIP_TRY(hipStreamWaitEvent(pl->compute, s.copied, 0));
    fxc_status st = enqueue_sector(p, (const uint8_t *)s.raw[0].p, (const uint8_t *)s.raw[1].p,
                                   first_frame, nf, *delay, (double *)s.acc.p, nullptr, pl->compute);
    if (st != FXC_OK) return st;
    HIP_TRY(hipMemcpyAsync(acc_out, s.acc.p, (size_t)kNumProducts * p->half * sizeof(double),
                           hipMemcpyDeviceToHost, pl->compute));
    HIP_TRY(hipMemcpyAsync(s.err, p->d_err.p, sizeof(int32_t), hipMemcpyDeviceToHost, pl->compute));
    HIP_TRY(hipEventRecord(s.done, pl->compute));
    s.ticket = t;
    pl->next_ticket = t + 1;
    *ticket = t;
    return FXC_OK;
}

extern "C" fxc_status fxc_pipeline_wait(fxc_pipeline *pl, int64_t ticket) {
    if (!pl || ticket < 0 || ticket >= pl->next_ticket) return fail(FXC_ERR_INVALID_ARG, "unknown ticket");
    auto &s = pl->slots[(size_t)(ticket % pl->depth)];
    // a slot reused since then: its later sector completes after this one
    HIP_TRY(hipEventSynchronize(s.done));
    if (*s.err != 0)
        return fail(FXC_ERR_HIP, "a sector kernel workgroup's bounded hand-over wait ran out "
                                 "(error word 0x%x); the spectra are not valid", (unsigned)*s.err);
    return FXC_OK;
}

extern "C" fxc_status fxc_pipeline_destroy(fxc_pipeline *pl) {
    pipeline_release(pl);
    return FXC_OK;
}
