// Device code of the FX hot path for gfx950 (MI355X).
//
// One workgroup owns a contiguous run of frames of one sector.  For every
// frame pair it: stages the packed words of both stations in LDS (shuffle
// applied once per word), decodes + integer-shifts straight into the first
// FFT pass's registers, runs an LDS Stockham complex FFT of N/2 points per
// station, untangles the real spectrum (realfft's method), and accumulates
// ACF/XCF products into FP64 registers owned per output bin.  Partials are
// written once per workgroup and reduced in a fixed order (deterministic).
#include "fxc_internal.h"

namespace fxc {

#define FXC_PI 3.14159265358979323846

// ------------------------------------------------------------------ complex
__device__ __forceinline__ float2 c_add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 c_sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 c_mul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 c_conj(float2 a) { return make_float2(a.x, -a.y); }
struct cd {
    double x, y;
};
__device__ __forceinline__ cd d_mul(cd a, cd b) { return cd{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }

// LDS index padding: one spare complex slot every 16 keeps the Stockham
// stride-E writes of the first pass free of bank conflicts.
__device__ __forceinline__ int pad(int i) { return i + (i >> 4); }

// ------------------------------------------------------------ delay model
// Rust `v as i64`: truncation with saturation, NaN -> 0.
__device__ __forceinline__ int64_t f64_to_i64_sat(double v) {
    if (v != v) return 0;
    if (v >= 9223372036854775807.0) return INT64_MAX;
    if (v <= -9223372036854775808.0) return INT64_MIN;
    return (int64_t)v;
}
__device__ __forceinline__ uint64_t f64_to_u64_sat(double v) {
    if (v != v || v <= 0.0) return 0;
    if (v >= 18446744073709551615.0) return UINT64_MAX;
    return (uint64_t)v;
}

// tau at the frame midpoint: main.rs:184-230.  No FP contraction so the
// result is bit-identical to the reference's operation order.
__device__ void frame_taus(int64_t idx, const DelayDev &c, double &tau1, double &tau2,
                           double &tf1, double &tf2) {
#pragma clang fp contract(off)
    double t_mid_local = ((double)idx + 0.5) * c.frame_dt;
    double t_mid = t_mid_local + c.time_offset_s;
    if (c.table != nullptr && c.table_len > 0) {
        double fl = floor(t_mid_local);
        uint64_t sec_idx = f64_to_u64_sat(fl > 0.0 ? fl : 0.0);
        uint64_t last = (uint64_t)c.table_len - 1;
        uint64_t i0 = sec_idx < last ? sec_idx : last;
        uint64_t i1 = (i0 + 1) < last ? (i0 + 1) : last;
        double frac;
        if (i1 == i0) {
            frac = 0.0;
        } else {
            frac = t_mid_local - (double)i0;
            if (frac < 0.0) frac = 0.0;
            if (frac > 1.0) frac = 1.0;
        }
        double gd_t = c.table[i0] * (1.0 - frac) + c.table[i1] * frac;
        double net = gd_t + c.coarse + c.user_samples / c.fs + c.extra_rate * t_mid;
        double c1 = c.c1d + c.c1r * t_mid;
        double c2 = c.c2d + c.c2r * t_mid;
        double rel = c2 - c1;
        double full = net + rel;
        tau1 = full - c.d_seek;
        tau2 = 0.0;
        tf1 = full;
        tf2 = 0.0;
    } else {
        double t1 = c.nd1 + c.tr1 * t_mid + 0.5 * c.ta1 * (t_mid * t_mid);
        double t2 = c.nd2 + c.tr2 * t_mid + 0.5 * c.ta2 * (t_mid * t_mid);
        tau1 = t1;
        tau2 = t2;
        tf1 = t1 + c.d_seek;
        tf2 = t2;
    }
}

// main.rs:126-130
__device__ __forceinline__ void split_delay(double d, double fs, int64_t &n, double &frac) {
#pragma clang fp contract(off)
    n = f64_to_i64_sat(round(d * fs));
    frac = d - ((double)n / fs);
}

// Phase -2*pi*f*tau reduced to [-pi, pi) in f64 (main.rs:110-120).
__device__ __forceinline__ double wrapped_carrier_phase(double f, double tau) {
#pragma clang fp contract(off)
    const double two_pi = 2.0 * FXC_PI;
    double phase = -two_pi * f * tau;
    double x = phase + FXC_PI;
    double r = fmod(x, two_pi);
    if (r < 0.0) r = r + fabs(two_pi);
    return r - FXC_PI;
}

// from_polar(1.0f32, theta as f32): the f32 cos/sin are taken as the
// correctly rounded values of the double-precision functions.
__device__ __forceinline__ float2 polar1_f32(float th) {
    return make_float2((float)cos((double)th), (float)sin((double)th));
}

__device__ __forceinline__ float2 carrier_phase_f32(double f, double tau) {
    if (f == 0.0) return make_float2(1.0f, 0.0f);
    return polar1_f32((float)wrapped_carrier_phase(f, tau));
}
__device__ __forceinline__ cd carrier_phase_f64(double f, double tau) {
    if (f == 0.0) return cd{1.0, 0.0};
    double w = wrapped_carrier_phase(f, tau);
    return cd{cos(w), sin(w)};
}

__device__ void compute_frame_delay(int64_t idx, const DelayDev &c, FrameDelayDev &o,
                                    cd *fr1_f64, cd *fr2_f64) {
    double tau1, tau2, tf1, tf2;
    frame_taus(idx, c, tau1, tau2, tf1, tf2);
    split_delay(tau1, c.fs, o.int1, o.frac1);
    split_delay(tau2, c.fs, o.int2, o.frac2);
    float2 a = carrier_phase_f32(c.lo1, tf1);
    float2 b = carrier_phase_f32(c.lo2, tf2);
    o.fr1[0] = a.x;
    o.fr1[1] = a.y;
    o.fr2[0] = b.x;
    o.fr2[1] = b.y;
    if (fr1_f64) *fr1_f64 = carrier_phase_f64(c.lo1, tf1);
    if (fr2_f64) *fr2_f64 = carrier_phase_f64(c.lo2, tf2);
}

// Per-frame values every thread of the workgroup needs, computed once per
// frame (one thread per frame) and broadcast through LDS.
struct FrameParam {
    int32_t s1, s2;      // integer shifts clamped to [-N, N]
    float2 frmix;        // fr_lo1 * conj(fr_lo2) in f32 (plain form)
    cd start;            // XCF phasor at overlap index 0
    cd pw[kMaxPow];      // XCF phase step^(2^b)
};

__device__ __forceinline__ int32_t clamp_shift(int64_t s, int n) {
    return (int32_t)(s > n ? n : (s < -n ? -n : s));
}

template <int PREC>
__device__ void make_frame_param(int64_t idx, const SectorArgs &a, int logm, FrameParam &p) {
    FrameDelayDev d;
    cd fr1, fr2;
    compute_frame_delay(idx, a.dly, d, &fr1, &fr2);
    p.s1 = clamp_shift(d.int1, a.N);
    p.s2 = clamp_shift(d.int2, a.N);
    cd step;
    {
#pragma clang fp contract(off)
        // main.rs:133-149 xcf_phase_start_and_step
        double two_pi_df = -2.0 * FXC_PI * a.df;
        double step1 = two_pi_df * d.frac1;
        double step2 = two_pi_df * d.frac2;
        double ang0 = step1 * (double)a.a1s - step2 * (double)a.a2s;
        double angs = step1 - step2;
        if (PREC == 0) {
            float2 f1 = make_float2(d.fr1[0], d.fr1[1]);
            float2 f2 = make_float2(d.fr2[0], d.fr2[1]);
            p.frmix = c_mul(f1, c_conj(f2));
            float2 st = polar1_f32((float)ang0);
            float2 sp = polar1_f32((float)angs);
            p.start = cd{(double)st.x, (double)st.y};
            step = cd{(double)sp.x, (double)sp.y};
        } else {
            cd mix = d_mul(fr1, cd{fr2.x, -fr2.y});
            p.frmix = make_float2(1.0f, 0.0f);
            p.start = d_mul(mix, cd{cos(ang0), sin(ang0)});
            step = cd{cos(angs), sin(angs)};
        }
    }
    p.pw[0] = step;
    for (int b = 1; b < logm; ++b) p.pw[b] = d_mul(p.pw[b - 1], p.pw[b - 1]);
}

// ------------------------------------------------------------- decode
__device__ __forceinline__ uint32_t shuffle_word(uint32_t w, const StationDev &st) {
    if (st.shuffle_kind == SHUF_PAIRSWAP) {
        w = ((w & 0xAAAAAAAAu) >> 1) | ((w & 0x55555555u) << 1);
    } else if (st.shuffle_kind == SHUF_GENERIC) {
        uint32_t o = 0;
#pragma unroll
        for (int i = 0; i < 32; ++i) o |= ((w >> st.shifts[i]) & 1u) << i;
        w = o;
    }
    return w;
}

// Stage the frame's packed words (little-endian, shuffle applied) in LDS.
template <int T>
__device__ __forceinline__ void stage_words(const uint8_t *frame, const StationDev &st,
                                            uint32_t *words) {
    const int t = threadIdx.x;
    if (st.aligned) {
        const uint32_t *src = reinterpret_cast<const uint32_t *>(frame);
        for (int w = t; w < st.nwords; w += T) words[w] = shuffle_word(__builtin_nontemporal_load(src + w), st);
    } else {
        for (int w = t; w < st.nwords; w += T) {
            uint32_t v = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                int b = 4 * w + j;
                if (b < st.bpf) v |= (uint32_t)frame[b] << (8 * j);
            }
            words[w] = shuffle_word(v, st);
        }
    }
}

// Level of a code; 1/2-bit levels come from registers, wider from LDS/global.
__device__ __forceinline__ float level_of(uint32_t code, const StationDev &st,
                                          const float *lev_lds) {
    if (st.bits == 2) {
        float lo = (code & 1u) ? st.lev4[1] : st.lev4[0];
        float hi = (code & 1u) ? st.lev4[3] : st.lev4[2];
        return (code & 2u) ? hi : lo;
    }
    if (st.bits == 1) return code ? st.lev4[1] : st.lev4[0];
    if (st.bits <= 8) return lev_lds[code];
    return st.levels[code];
}

// Decoded, integer-shifted sample n of the frame (utils.rs:215-376 then
// utils.rs:126-145): sample n holds decoded sample n - shift, zero outside.
// The LSB->USB flip negates odd decode-order samples.
__device__ __forceinline__ float shifted_sample(int n, int shift, int N, const uint32_t *words,
                                                const StationDev &st, const float *lev_lds) {
    int src = n - shift;
    if ((unsigned)src >= (unsigned)N) return 0.0f;
    uint32_t bitpos = (uint32_t)src * (uint32_t)st.bits;
    uint32_t w0 = bitpos >> 5, off = bitpos & 31u;
    uint32_t v = words[w0] >> off;
    if (off + (uint32_t)st.bits > 32u) v |= words[w0 + 1] << (32u - off);
    float val = level_of(v & st.code_mask, st, lev_lds);
    return (st.lsb && (src & 1)) ? -val : val;
}

// ------------------------------------------------------------------ FFT
__device__ __forceinline__ void dft2(float2 &a, float2 &b) {
    float2 t = a;
    a = c_add(t, b);
    b = c_sub(t, b);
}
__device__ __forceinline__ void dft4(float2 &a0, float2 &a1, float2 &a2, float2 &a3) {
    float2 s02 = c_add(a0, a2), d02 = c_sub(a0, a2);
    float2 s13 = c_add(a1, a3), d13 = c_sub(a1, a3);
    a0 = c_add(s02, s13);
    a2 = c_sub(s02, s13);
    a1 = make_float2(d02.x + d13.y, d02.y - d13.x);
    a3 = make_float2(d02.x - d13.y, d02.y + d13.x);
}
__device__ __forceinline__ void dft8(float2 *a) {
    float2 e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
    float2 o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
    dft4(e0, e1, e2, e3);
    dft4(o0, o1, o2, o3);
    const float r = 0.70710678118654752440f;
    float2 t1 = make_float2((o1.x + o1.y) * r, (o1.y - o1.x) * r);
    float2 t2 = make_float2(o2.y, -o2.x);
    float2 t3 = make_float2((o3.y - o3.x) * r, (-o3.x - o3.y) * r);
    a[0] = c_add(e0, o0);
    a[4] = c_sub(e0, o0);
    a[1] = c_add(e1, t1);
    a[5] = c_sub(e1, t1);
    a[2] = c_add(e2, t2);
    a[6] = c_sub(e2, t2);
    a[3] = c_add(e3, t3);
    a[7] = c_sub(e3, t3);
}
template <int R>
__device__ __forceinline__ void dftR(float2 *a) {
    if constexpr (R == 2) dft2(a[0], a[1]);
    else if constexpr (R == 4) dft4(a[0], a[1], a[2], a[3]);
    else dft8(a);
}

// One Stockham radix-R pass (Govindaraju et al. 2008 formulation) over two
// in-place LDS buffers: butterfly j reads z[j + r*M/R], twiddles by
// W_M^{r*(j%NS)*M/(NS*R)}, writes z[(j/NS)*NS*R + j%NS + r*NS].
template <int E, int T, int NS>
__device__ __forceinline__ void stockham_rest(float2 *za, float2 *zb, const float2 *tw) {
    constexpr int M = E * T;
    if constexpr (NS < M) {
        constexpr int R = (M / NS < E) ? (M / NS) : E;
        constexpr int BPT = E / R;
        const int t = threadIdx.x;
        float2 a[BPT][R], b[BPT][R];
#pragma unroll
        for (int u = 0; u < BPT; ++u) {
            const int j = t + T * u;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                a[u][r] = za[pad(j + r * (M / R))];
                b[u][r] = zb[pad(j + r * (M / R))];
            }
            if constexpr (NS > 1) {
                const int jm = j % NS;
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    float2 w = tw[r * jm * (M / (NS * R))];
                    a[u][r] = c_mul(a[u][r], w);
                    b[u][r] = c_mul(b[u][r], w);
                }
            }
            dftR<R>(a[u]);
            dftR<R>(b[u]);
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < BPT; ++u) {
            const int j = t + T * u;
            const int base = (j / NS) * NS * R + (j % NS);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                za[pad(base + r * NS)] = a[u][r];
                zb[pad(base + r * NS)] = b[u][r];
            }
        }
        __syncthreads();
        stockham_rest<E, T, NS * R>(za, zb, tw);
    }
}

// First pass from registers: v[r] = z[t + T*r] (NS = 1, no twiddles).
template <int E, int T>
__device__ __forceinline__ void stockham_first(float2 *za, float2 *zb, float2 *va, float2 *vb) {
    const int t = threadIdx.x;
    dftR<E>(va);
    dftR<E>(vb);
#pragma unroll
    for (int r = 0; r < E; ++r) {
        za[pad(t * E + r)] = va[r];
        zb[pad(t * E + r)] = vb[r];
    }
    __syncthreads();
}

// X[k] of the real N-point input from Z = FFT_{N/2}(x[2m] + i x[2m+1])
// (realfft's split method): X = E + W_N^k O, E = (Z_k + conj Z_{M-k})/2,
// O = (Z_k - conj Z_{M-k})/(2i).  DC and Nyquist are purely real.
__device__ __forceinline__ float2 untangle(const float2 *zs, int k, int M, const float2 *tw_post) {
    if (k == 0 || k == M) {
        float2 z0 = zs[0];
        return make_float2(k == 0 ? z0.x + z0.y : z0.x - z0.y, 0.0f);
    }
    float2 A = zs[pad(k)];
    float2 Bc = zs[pad(M - k)];
    float ex = 0.5f * (A.x + Bc.x), ey = 0.5f * (A.y - Bc.y);
    float dx = 0.5f * (A.x - Bc.x), dy = 0.5f * (A.y + Bc.y);
    float2 w = tw_post[k];
    return make_float2(ex + (w.x * dy - w.y * dx), ey - (w.x * dx + w.y * dy));
}

__device__ __forceinline__ float norm_sqr(float2 a) { return a.x * a.x + a.y * a.y; }

// ------------------------------------------------------- fused kernel
template <int E, int T, int PREC>
__global__ __launch_bounds__(T) void k_sector(SectorArgs a) {
    constexpr int M = E * T;
    constexpr int MP = M + M / 16 + 2;
    constexpr int NB = E + 1;
    constexpr int LOGT = __builtin_ctz(T);
    constexpr int LOGM = __builtin_ctz(M);
    static_assert(LOGM <= kMaxPow, "phase power table too small");
    __shared__ float2 zbuf[2][MP];
    __shared__ FrameParam fp[kFrameBatch];
    __shared__ float lev[2][256];

    const int t = threadIdx.x;
    const int N = 2 * M;
    const int64_t f_begin = (int64_t)blockIdx.x * a.frames_per_wg;
    int64_t f_end = f_begin + a.frames_per_wg;
    if (f_end > a.nf) f_end = a.nf;

#pragma unroll
    for (int s = 0; s < 2; ++s)
        if (a.st[s].bits > 2 && a.st[s].bits <= 8)
            for (int i = t; i < 256; i += T) lev[s][i] = (i <= (int)a.st[s].code_mask) ? a.st[s].levels[i] : 0.0f;

    double acc[4][NB];
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int i = 0; i < NB; ++i) acc[p][i] = 0.0;

    const bool do_acf = (a.flags & FLAG_ACF) != 0;
    const bool do_xcf = (a.flags & FLAG_XCF) != 0;
    const bool ov_only = (a.flags & FLAG_ACF_OVERLAP) != 0;
    const int gs = (a.grid == 0) ? a.a1s : a.a2s;   // grid index of overlap bin 0
    const int ge = gs + a.ov;
    const int sb_other = (a.grid == 0) ? (a.a2s - a.a1s) : (a.a1s - a.a2s);
    float2 *z0 = zbuf[0];
    float2 *z1 = zbuf[1];
    uint32_t *w0 = reinterpret_cast<uint32_t *>(zbuf[0]);
    uint32_t *w1 = reinterpret_cast<uint32_t *>(zbuf[1]);
    __syncthreads();

    for (int64_t fb = f_begin; fb < f_end; fb += kFrameBatch) {
        const int nb = (int)((f_end - fb) < kFrameBatch ? (f_end - fb) : kFrameBatch);
        if (t < nb) make_frame_param<PREC>(a.first_frame + fb + t, a, LOGM, fp[t]);
        __syncthreads();
        for (int jf = 0; jf < nb; ++jf) {
            const int64_t f = fb + jf;
            stage_words<T>(a.raw1 + f * a.st[0].bpf, a.st[0], w0);
            stage_words<T>(a.raw2 + f * a.st[1].bpf, a.st[1], w1);
            __syncthreads();
            const int s1 = fp[jf].s1, s2 = fp[jf].s2;
            float2 va[E], vb[E];
#pragma unroll
            for (int r = 0; r < E; ++r) {
                const int m = t + T * r;
                va[r] = make_float2(shifted_sample(2 * m, s1, N, w0, a.st[0], lev[0]),
                                    shifted_sample(2 * m + 1, s1, N, w0, a.st[0], lev[0]));
                vb[r] = make_float2(shifted_sample(2 * m, s2, N, w1, a.st[1], lev[1]),
                                    shifted_sample(2 * m + 1, s2, N, w1, a.st[1], lev[1]));
            }
            __syncthreads();
            stockham_first<E, T>(z0, z1, va, vb);
            stockham_rest<E, T, E>(z0, z1, a.tw_fft);

            // ---- products on the output grid (main.rs:2491-2568)
            const float2 frmix = fp[jf].frmix;
            bool have = false;
            cd P = cd{0.0, 0.0};
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                const int g = t + T * i;
                if (i < E || g <= M) {
                const bool in_ov = (g >= gs) && (g < ge);
                float2 Xg = untangle(a.grid == 0 ? z0 : z1, g, M, a.tw_post);
                float pg = norm_sqr(Xg);
                float2 Xo = make_float2(0.0f, 0.0f);
                if (in_ov) Xo = untangle(a.grid == 0 ? z1 : z0, g + sb_other, M, a.tw_post);
                if (do_acf) {
                    // grid station: every bin (or overlap only); other station: overlap
                    const float pgrid = (!ov_only || in_ov) ? pg : 0.0f;
                    const float poth = in_ov ? norm_sqr(Xo) : 0.0f;
                    acc[0][i] += (double)(a.grid == 0 ? pgrid : poth);
                    acc[1][i] += (double)(a.grid == 0 ? poth : pgrid);
                }
                if (in_ov) {
                    if (do_xcf) {
                        const int k = g - gs;
                        if (!have) {
                            P = fp[jf].start;
#pragma unroll
                            for (int b = 0; b < LOGM; ++b)
                                if ((k >> b) & 1) P = d_mul(P, fp[jf].pw[b]);
                            have = true;
                        } else {
                            P = d_mul(P, fp[jf].pw[LOGT]);
                        }
                        const float2 ph = make_float2((float)P.x, (float)P.y);
                        const float2 X1 = (a.grid == 0) ? Xg : Xo;
                        const float2 X2 = (a.grid == 0) ? Xo : Xg;
                        float2 v = c_mul(X1, c_conj(X2));
                        if (PREC == 0) v = c_mul(v, frmix);
                        v = c_mul(v, ph);
                        acc[2][i] += (double)v.x;
                        acc[3][i] += (double)v.y;
                    }
                }
                }
            }
            __syncthreads();
        }
    }

    // ---- workgroup partials [wg][ph|11|22|12re|12im][half]
    const int half = M + 1;
    double *out = a.part + (size_t)blockIdx.x * kNumProducts * half;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int g = t + T * i;
        if (g <= M) {
            out[P_PH * half + g] = 0.0;
            out[P_11 * half + g] = acc[0][i];
            out[P_22 * half + g] = acc[1][i];
            out[P_12R * half + g] = acc[2][i];
            out[P_12I * half + g] = acc[3][i];
        }
    }
}

// ---------------------------------------------------- deterministic reduce
__global__ void k_reduce_chunks(const double *part, int n_wg, int cols, int chunk,
                                double *scratch) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (col >= cols) return;
    const int w_begin = c * chunk;
    int w_end = w_begin + chunk;
    if (w_end > n_wg) w_end = n_wg;
    double s = 0.0;
    for (int w = w_begin; w < w_end; ++w) s += part[(size_t)w * cols + col];
    scratch[(size_t)c * cols + col] = s;
}

__global__ void k_reduce_final(const double *scratch, int n_chunks, int cols, double *out) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= cols) return;
    double s = 0.0;
    for (int c = 0; c < n_chunks; ++c) s += scratch[(size_t)c * cols + col];
    out[col] = s;
}

hipError_t launch_reduce(const double *part, int n_wg, int half, double *out, double *scratch,
                         int n_chunks, hipStream_t stream) {
    const int cols = kNumProducts * half;
    if (n_chunks < 1) n_chunks = 1;
    if (n_chunks > n_wg) n_chunks = n_wg;
    const int chunk = (n_wg + n_chunks - 1) / n_chunks;
    n_chunks = (n_wg + chunk - 1) / chunk;
    dim3 g1((cols + 255) / 256, n_chunks);
    hipLaunchKernelGGL(k_reduce_chunks, g1, dim3(256), 0, stream, part, n_wg, cols, chunk, scratch);
    hipLaunchKernelGGL(k_reduce_final, dim3((cols + 255) / 256), dim3(256), 0, stream, scratch,
                       n_chunks, cols, out);
    return hipGetLastError();
}

// ---------------------------------------------------------- dispatch
// Elements per thread E and threads per frame T for a given N (M = N/2 = E*T).
static void shape_for(int N, int &E, int &T) {
    const int M = N / 2;
    if (M >= 512) {
        E = 8;
        T = M / 8;
    } else if (M == 256) {
        E = 4;
        T = 64;
    } else {
        E = 2;
        T = M / 2;
    }
}

bool sector_supported(int N) { return N >= 64 && N <= 16384 && (N & (N - 1)) == 0; }
int sector_threads_per_frame(int N) {
    int E, T;
    shape_for(N, E, T);
    return T;
}

template <int E, int T>
static hipError_t launch_sector_t(const SectorArgs &a, int precision, int n_wg, hipStream_t s) {
    if (precision == 0)
        hipLaunchKernelGGL((k_sector<E, T, 0>), dim3(n_wg), dim3(T), 0, s, a);
    else
        hipLaunchKernelGGL((k_sector<E, T, 1>), dim3(n_wg), dim3(T), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_sector(const SectorArgs &a, int precision, int n_wg, hipStream_t s) {
    switch (a.N) {
    case 64: return launch_sector_t<2, 16>(a, precision, n_wg, s);
    case 128: return launch_sector_t<2, 32>(a, precision, n_wg, s);
    case 256: return launch_sector_t<2, 64>(a, precision, n_wg, s);
    case 512: return launch_sector_t<4, 64>(a, precision, n_wg, s);
    case 1024: return launch_sector_t<8, 64>(a, precision, n_wg, s);
    case 2048: return launch_sector_t<8, 128>(a, precision, n_wg, s);
    case 4096: return launch_sector_t<8, 256>(a, precision, n_wg, s);
    case 8192: return launch_sector_t<8, 512>(a, precision, n_wg, s);
    case 16384: return launch_sector_t<8, 1024>(a, precision, n_wg, s);
    default: return hipErrorInvalidValue;
    }
}

// ------------------------------------------ per-function entry kernels
// Decode (+ shift) of one frame per workgroup into out[f*N + n].
template <int T>
__global__ __launch_bounds__(T) void k_decode_frames(SectorArgs a, int station,
                                                     const int64_t *shifts, float *out) {
    __shared__ uint32_t words[16384 / 2 + 8];
    __shared__ float lev[256];
    const StationDev &st = a.st[station];
    const int t = threadIdx.x;
    const int64_t f = blockIdx.x;
    if (st.bits > 2 && st.bits <= 8)
        for (int i = t; i < 256; i += T) lev[i] = (i <= (int)st.code_mask) ? st.levels[i] : 0.0f;
    const uint8_t *raw = (station == 0 ? a.raw1 : a.raw2) + f * st.bpf;
    stage_words<T>(raw, st, words);
    __syncthreads();
    const int s = shifts ? clamp_shift(shifts[f], a.N) : 0;
    for (int n = t; n < a.N; n += T) out[f * a.N + n] = shifted_sample(n, s, a.N, words, st, lev);
}

hipError_t launch_decode_frames(const SectorArgs &a, int station, const int64_t *shifts,
                                float *out, hipStream_t s) {
    if (a.nf <= 0) return hipSuccess;
    if (a.st[station].nwords > 16384 / 2 + 8) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_decode_frames<256>, dim3((unsigned)a.nf), dim3(256), 0, s, a, station,
                       shifts, out);
    return hipGetLastError();
}

// Forward R2C of two frames per workgroup with the fused kernel's FFT code.
template <int E, int T>
__global__ __launch_bounds__(T) void k_rfft_frames(SectorArgs a, const float *in, int64_t nframes,
                                                   float2 *out) {
    constexpr int M = E * T;
    constexpr int MP = M + M / 16 + 2;
    __shared__ float2 zbuf[2][MP];
    const int t = threadIdx.x;
    const int64_t fa = 2 * (int64_t)blockIdx.x, fbk = fa + 1;
    const bool hb = fbk < nframes;
    float2 va[E], vb[E];
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const int m = t + T * r;
        va[r] = reinterpret_cast<const float2 *>(in + fa * 2 * M)[m];
        vb[r] = hb ? reinterpret_cast<const float2 *>(in + fbk * 2 * M)[m] : make_float2(0.f, 0.f);
    }
    stockham_first<E, T>(zbuf[0], zbuf[1], va, vb);
    stockham_rest<E, T, E>(zbuf[0], zbuf[1], a.tw_fft);
    for (int k = t; k <= M; k += T) {
        out[fa * (M + 1) + k] = untangle(zbuf[0], k, M, a.tw_post);
        if (hb) out[fbk * (M + 1) + k] = untangle(zbuf[1], k, M, a.tw_post);
    }
}

// Inverse C2R (scaled 1/N): build Z' = E' + i O' from the half spectrum,
// IFFT via conj(FFT(conj(.))), interleave, scale.  DC/Nyquist imaginary
// parts are ignored (the reference zeroes them first).
template <int E, int T>
__device__ __forceinline__ void irfft_prepare(const float2 *X, int M, const float2 *tw_post,
                                              float2 *v) {
    const int t = threadIdx.x;
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const int k = t + T * r;
        float2 z;
        if (k == 0) {
            z = make_float2(X[0].x + X[M].x, X[0].x - X[M].x);
        } else {
            float2 A = X[k], Bc = X[M - k];
            float er = A.x + Bc.x, ei = A.y - Bc.y;
            float dr = A.x - Bc.x, di = A.y + Bc.y;
            float2 w = tw_post[k];
            float orr = dr * w.x - di * w.y;
            float oi = dr * w.y + di * w.x;
            z = make_float2(er - oi, ei + orr);
        }
        v[r] = c_conj(z);
    }
}

template <int E, int T>
__global__ __launch_bounds__(T) void k_irfft_frames(SectorArgs a, const float2 *in,
                                                    int64_t nframes, float *out) {
    constexpr int M = E * T;
    constexpr int MP = M + M / 16 + 2;
    __shared__ float2 zbuf[2][MP];
    const int t = threadIdx.x;
    const int64_t fa = 2 * (int64_t)blockIdx.x, fbk = fa + 1;
    const bool hb = fbk < nframes;
    float2 va[E], vb[E];
    irfft_prepare<E, T>(in + fa * (M + 1), M, a.tw_post, va);
    if (hb) irfft_prepare<E, T>(in + fbk * (M + 1), M, a.tw_post, vb);
    else
        for (int r = 0; r < E; ++r) vb[r] = make_float2(0.f, 0.f);
    stockham_first<E, T>(zbuf[0], zbuf[1], va, vb);
    stockham_rest<E, T, E>(zbuf[0], zbuf[1], a.tw_fft);
    const float scale = 1.0f / (float)(2 * M);
    for (int m = t; m < M; m += T) {
        float2 za = zbuf[0][pad(m)];
        out[fa * 2 * M + 2 * m] = za.x * scale;
        out[fa * 2 * M + 2 * m + 1] = -za.y * scale;
        if (hb) {
            float2 zb = zbuf[1][pad(m)];
            out[fbk * 2 * M + 2 * m] = zb.x * scale;
            out[fbk * 2 * M + 2 * m + 1] = -zb.y * scale;
        }
    }
}

template <int E, int T>
static hipError_t launch_fft_t(const SectorArgs &a, const void *in, int64_t nframes, void *out,
                               bool inverse, hipStream_t s) {
    dim3 grid((unsigned)((nframes + 1) / 2));
    if (inverse)
        hipLaunchKernelGGL((k_irfft_frames<E, T>), grid, dim3(T), 0, s, a, (const float2 *)in,
                           nframes, (float *)out);
    else
        hipLaunchKernelGGL((k_rfft_frames<E, T>), grid, dim3(T), 0, s, a, (const float *)in,
                           nframes, (float2 *)out);
    return hipGetLastError();
}

static hipError_t launch_fft(const SectorArgs &a, const void *in, int64_t nframes, void *out,
                             bool inverse, hipStream_t s) {
    if (nframes <= 0) return hipSuccess;
    switch (a.N) {
    case 64: return launch_fft_t<2, 16>(a, in, nframes, out, inverse, s);
    case 128: return launch_fft_t<2, 32>(a, in, nframes, out, inverse, s);
    case 256: return launch_fft_t<2, 64>(a, in, nframes, out, inverse, s);
    case 512: return launch_fft_t<4, 64>(a, in, nframes, out, inverse, s);
    case 1024: return launch_fft_t<8, 64>(a, in, nframes, out, inverse, s);
    case 2048: return launch_fft_t<8, 128>(a, in, nframes, out, inverse, s);
    case 4096: return launch_fft_t<8, 256>(a, in, nframes, out, inverse, s);
    case 8192: return launch_fft_t<8, 512>(a, in, nframes, out, inverse, s);
    case 16384: return launch_fft_t<8, 1024>(a, in, nframes, out, inverse, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_rfft_frames(const SectorArgs &a, const float *in, int64_t nframes, float2 *out,
                              hipStream_t s) {
    return launch_fft(a, in, nframes, out, false, s);
}
hipError_t launch_irfft_frames(const SectorArgs &a, const float2 *in, int64_t nframes, float *out,
                               hipStream_t s) {
    return launch_fft(a, in, nframes, out, true, s);
}

// Delay entries, one thread per frame.
__global__ void k_delay_eval(DelayDev d, int64_t first, int64_t n, FrameDelayDev *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    FrameDelayDev o;
    compute_frame_delay(first + i, d, o, nullptr, nullptr);
    out[i] = o;
}

hipError_t launch_delay_eval(const DelayDev &d, int64_t first, int64_t n, FrameDelayDev *out,
                             hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_delay_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d, first,
                       n, out);
    return hipGetLastError();
}

// ----------------------------------------------------------- quantise
// utils.rs:378-410: nearest level (first minimum of |x - level| over the
// levels in order), codes packed LSB-first into 32-bit words, each word put
// through the output bit map (out bit shuffle[pos] <- in bit pos).
__device__ __forceinline__ uint32_t quant_code(float x, const float *lv, int nl) {
    uint32_t best = 0;
    float min_err = 3.40282347e+38f;
    for (int k = 0; k < nl; ++k) {
        float err = fabsf(x - lv[k]);
        if (err < min_err) {
            min_err = err;
            best = (uint32_t)k;
        }
    }
    return best;
}

__device__ __forceinline__ uint32_t shuffle_out_word(uint32_t w, const int32_t *sh) {
    uint32_t o = 0;
#pragma unroll
    for (int p = 0; p < 32; ++p) o |= ((w >> p) & 1u) << sh[p];
    return o;
}

// One thread per output word of one frame.
__device__ __forceinline__ uint32_t quant_word(const float *x, int N, int w, int bits,
                                               const float *lv, int nl, uint32_t mask) {
    uint64_t acc = 0;
    const int64_t b0 = (int64_t)w * 32, b1 = b0 + 32;
    int64_t s = b0 / bits;
    for (; s * bits < b1 && s < N; ++s) {
        uint64_t code = (uint64_t)(quant_code(x[s], lv, nl) & mask);
        int64_t pos = s * bits - b0;  // may be negative for the straddling first sample
        if (pos >= 0) acc |= code << pos;
        else acc |= code >> (-pos);
    }
    return (uint32_t)(acc & 0xFFFFFFFFull);
}

__global__ void k_quantise_frames(SectorArgs a, const float *in, int64_t nframes, uint8_t *out) {
    __shared__ float lv[256];
    const int nl = a.out_nlevels;
    for (int i = threadIdx.x; i < nl && i < 256; i += blockDim.x) lv[i] = a.out_levels[i];
    __syncthreads();
    const int nwo = a.bpf_out / 4;
    const uint32_t mask = (a.bit_out >= 32) ? 0xFFFFFFFFu : ((1u << a.bit_out) - 1u);
    const int64_t f = blockIdx.x;
    for (int w = threadIdx.x; w < nwo; w += blockDim.x) {
        uint32_t word = quant_word(in + f * a.N, a.N, w, a.bit_out, lv, nl < 256 ? nl : 256, mask);
        uint32_t sh = shuffle_out_word(word, a.out_shifts);
        reinterpret_cast<uint32_t *>(out + f * a.bpf_out)[w] = sh;
    }
}

hipError_t launch_quantise_frames(const SectorArgs &a, const float *in, int64_t nframes,
                                  uint8_t *out, hipStream_t s) {
    if (nframes <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_quantise_frames, dim3((unsigned)nframes), dim3(256), 0, s, a, in, nframes,
                       out);
    return hipGetLastError();
}

}  // namespace fxc
