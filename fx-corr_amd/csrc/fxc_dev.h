// Device-side building blocks shared by the fused kernels: complex and
// FP64-delay-model helpers, per-frame parameters, decode, radix-2/4/8/16
// butterflies, the LDS Stockham FFT and the realfft-style untangle.
#pragma once

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

// The value v, opaque to the optimizer from here on.  A frame loop that
// takes its thread indices through this forms the LDS addresses and
// twiddle loads derived from them inside the loop; otherwise the compiler
// hoists all of them out of the loop and keeps them in registers for the
// whole kernel, which doubles the Stockham kernels' register count and
// halves their occupancy (216 -> 120 VGPRs at N = 1024).
__device__ __forceinline__ int loop_local(int v) {
    asm volatile("" : "+v"(v));
    return v;
}

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
// x reduced to [-1/2, 1/2] turns
__device__ __forceinline__ double wrap_turns(double x) { return x - rint(x); }

// exp(2 pi i x) for x in turns (reduced mod 1 in FP64, f32 sincos)
__device__ __forceinline__ float2 phasor_turns(double x) {
    const float r = (float)wrap_turns(x);
    return make_float2(__builtin_amdgcn_cosf(r), __builtin_amdgcn_sinf(r));
}

__device__ __forceinline__ float2 c_scale(float2 a, float w) { return make_float2(a.x * w, a.y * w); }

struct FrameParam {
    int32_t s1, s2;      // integer shifts clamped to [-N, N]
    float2 frmix;        // fr_lo1 * conj(fr_lo2) in f32 (plain form)
    cd start;            // XCF phasor at overlap index 0 (plain form)
    cd pw[kMaxPow];      // XCF phase step^(2^b) (plain form)
    double phw, dth;     // chosen form: XCF phase at overlap index 0 (fringe
                         // difference included) and its step per index, turns
};

// Per-frame parameters of the phased-array sum (k_phased_params).
struct PhasedParam {
    int32_t s1, s2;        // integer shifts clamped to [-N, N]
    float2 fr1, fr2;       // fringe phasors (plain: main.rs:235-240, f32; chosen: FP64 phase)
    double th1, th2;       // chosen: rotation per bin, turns (-df frac)
};

__device__ __forceinline__ int32_t clamp_shift(int64_t s, int n) {
    return (int32_t)(s > n ? n : (s < -n ? -n : s));
}

template <int PREC>
__device__ void make_frame_param(int64_t idx, const SectorArgs &a, int logm, FrameParam &p) {
    FrameDelayDev d;
    compute_frame_delay(idx, a.dly, d, nullptr, nullptr);
    p.s1 = clamp_shift(d.int1, a.N);
    p.s2 = clamp_shift(d.int2, a.N);
    if (PREC != 0) {
        // chosen form: main.rs:133-149 and the fringe difference restated in
        // FP64 turns: w1 - w2 (-lo tau reduced mod 1) plus
        // -df (frac1 (a1s + k) - frac2 (a2s + k)) at overlap index k
        double tau1, tau2, tf1, tf2;
        frame_taus(idx, a.dly, tau1, tau2, tf1, tf2);
        const double w1 = a.dly.lo1 == 0.0 ? 0.0 : wrap_turns(-a.dly.lo1 * tf1);
        const double w2 = a.dly.lo2 == 0.0 ? 0.0 : wrap_turns(-a.dly.lo2 * tf2);
        p.frmix = make_float2(1.0f, 0.0f);
        p.start = cd{1.0, 0.0};
        p.dth = -a.df * (d.frac1 - d.frac2);
        p.phw = wrap_turns(w1 - w2 - a.df * (d.frac1 * (double)a.a1s - d.frac2 * (double)a.a2s));
        return;
    }
    cd step;
    {
#pragma clang fp contract(off)
        // main.rs:133-149 xcf_phase_start_and_step
        double two_pi_df = -2.0 * FXC_PI * a.df;
        double step1 = two_pi_df * d.frac1;
        double step2 = two_pi_df * d.frac2;
        double ang0 = step1 * (double)a.a1s - step2 * (double)a.a2s;
        double angs = step1 - step2;
        float2 f1 = make_float2(d.fr1[0], d.fr1[1]);
        float2 f2 = make_float2(d.fr2[0], d.fr2[1]);
        p.frmix = c_mul(f1, c_conj(f2));
        float2 st = polar1_f32((float)ang0);
        float2 sp = polar1_f32((float)angs);
        p.start = cd{(double)st.x, (double)st.y};
        step = cd{(double)sp.x, (double)sp.y};
    }
    p.phw = 0.0;
    p.dth = 0.0;
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

// Stage the frame's packed words (little-endian, shuffle applied) in LDS;
// T threads, t = this thread's index among them.
template <int T>
__device__ __forceinline__ void stage_words(const uint8_t *frame, const StationDev &st,
                                            uint32_t *words, int t) {
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
// Radix-16 as 4 x 4: y[n2][k1] = DFT4_{n1}(a[4 n1 + n2]) * W16^{n2 k1},
// X[k1 + 4 k2] = DFT4_{n2}(y[n2][k1]).
__device__ __forceinline__ void dft16(float2 *a) {
    float2 y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        y[n2][0] = a[n2];
        y[n2][1] = a[4 + n2];
        y[n2][2] = a[8 + n2];
        y[n2][3] = a[12 + n2];
        dft4(y[n2][0], y[n2][1], y[n2][2], y[n2][3]);
    }
    const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f;
    const float r2 = 0.70710678118654752440f;
    // W16^1 = (c1, -s1), W16^2 = (r2, -r2), W16^3 = (s1, -c1), W16^4 = -i,
    // W16^6 = (-r2, -r2), W16^9 = (-c1, s1)
    y[1][1] = c_mul(y[1][1], make_float2(c1, -s1));
    y[1][2] = c_mul(y[1][2], make_float2(r2, -r2));
    y[1][3] = c_mul(y[1][3], make_float2(s1, -c1));
    y[2][1] = c_mul(y[2][1], make_float2(r2, -r2));
    y[2][2] = make_float2(y[2][2].y, -y[2][2].x);
    y[2][3] = c_mul(y[2][3], make_float2(-r2, -r2));
    y[3][1] = c_mul(y[3][1], make_float2(s1, -c1));
    y[3][2] = c_mul(y[3][2], make_float2(-r2, -r2));
    y[3][3] = c_mul(y[3][3], make_float2(-c1, s1));
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        float2 b0 = y[0][k1], b1 = y[1][k1], b2 = y[2][k1], b3 = y[3][k1];
        dft4(b0, b1, b2, b3);
        a[k1] = b0;
        a[k1 + 4] = b1;
        a[k1 + 8] = b2;
        a[k1 + 12] = b3;
    }
}

template <int R>
__device__ __forceinline__ void dftR(float2 *a) {
    if constexpr (R == 2) dft2(a[0], a[1]);
    else if constexpr (R == 4) dft4(a[0], a[1], a[2], a[3]);
    else if constexpr (R == 8) dft8(a);
    else dft16(a);
}

// One Stockham radix-R pass (Govindaraju et al. 2008 formulation) over an
// in-place LDS buffer of M = E*T points, T threads per buffer (t = thread
// index within its buffer): butterfly j reads z[j + r*M/R], twiddles by
// W_M^{r*(j%NS)*M/(NS*R)}, writes z[(j/NS)*NS*R + j%NS + r*NS].  Every
// thread of the workgroup runs the same pass sequence (the barriers are
// workgroup-wide; each half of the workgroup owns one buffer).
template <int E, int T, int NS>
__device__ __forceinline__ void stockham_rest(float2 *z, const float2 *tw, int t) {
    constexpr int M = E * T;
    if constexpr (NS < M) {
        constexpr int R = (M / NS < E) ? (M / NS) : E;
        constexpr int BPT = E / R;
        float2 a[BPT][R];
#pragma unroll
        for (int u = 0; u < BPT; ++u) {
            const int j = t + T * u;
#pragma unroll
            for (int r = 0; r < R; ++r) a[u][r] = z[pad(j + r * (M / R))];
            if constexpr (NS > 1) {
                const int jm = j % NS;
#pragma unroll
                for (int r = 1; r < R; ++r) a[u][r] = c_mul(a[u][r], tw[r * jm * (M / (NS * R))]);
            }
            dftR<R>(a[u]);
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < BPT; ++u) {
            const int j = t + T * u;
            const int base = (j / NS) * NS * R + (j % NS);
#pragma unroll
            for (int r = 0; r < R; ++r) z[pad(base + r * NS)] = a[u][r];
        }
        __syncthreads();
        stockham_rest<E, T, NS * R>(z, tw, t);
    }
}

// First pass from registers: v[r] = z[t + T*r] (NS = 1, no twiddles).
template <int E, int T>
__device__ __forceinline__ void stockham_first(float2 *z, float2 *v, int t) {
    dftR<E>(v);
#pragma unroll
    for (int r = 0; r < E; ++r) z[pad(t * E + r)] = v[r];
    __syncthreads();
}

// ------------------------------------------- FP16 MFMA DFT-32 (chosen form)
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

// One 16-row tile of a DFT-32 GEMM, output column tile t (t&1: k1 half,
// t>>1: re|im): y[r] = Y[row 4(l/16) + r][k1 = l%16 + 16(t&1)] (re or im).
// A operands a0/a1 (K-steps 0/1) hold, for lane l, row l%16 and the re|im
// pairs of m1 = 16 kap + l/16 + 4q (q = 0..3); B the matching hi / lo FP16
// parts of the DFT-32 coefficients.  The lo products (FP16 subnormals
// included) go first into the one FP32 accumulator, the hi products after.
__device__ __forceinline__ f4 mf_stage1(half8 a0, half8 a1, half8 bh0, half8 bl0, half8 bh1,
                                        half8 bl1) {
    f4 d = {0.f, 0.f, 0.f, 0.f};
    d = __builtin_amdgcn_mfma_f32_16x16x32_f16(a0, bl0, d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1, bl1, d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_16x16x32_f16(a0, bh0, d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1, bh1, d, 0, 0, 0);
    return d;
}

// The first Stockham pass (NS = 1, radix 32) of this half's M-point buffer
// as FP16 MFMA GEMMs: z[m2*32 + k1] = sum_m1 W32^{m1 k1} z[m2 + M2 m1],
// M2 = M/32 rows of 32-point DFTs in 16-row tiles, E/8 tiles per wave.  The
// inputs are decoded levels, exact in FP16 (the plan checks this), so the
// conversion is exact.  b1: the hi/lo operand table in global memory
// (16 x 64 uint4, [(kap*4 + t)*2 + lohi][lane]).  Workgroup-wide barriers:
// every thread of the workgroup calls this.
template <int E, int T>
__device__ __forceinline__ void stockham_first_mfma(float2 *z, const uint4 *b1, int t) {
    constexpr int M = E * T, M2 = M / 32, TPW = E / 8;
    static_assert(T % 64 == 0 && TPW >= 1 && TPW * (T / 64) * 16 == M2, "tile layout");
    const int l = t & 63;
    const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
    const int c = l & 15, g = l >> 4;
    half8 A[TPW][2];
#pragma unroll
    for (int tp = 0; tp < TPW; ++tp) {
        const int m2 = 16 * (wv * TPW + tp) + c;
        uint32_t d[8];
#pragma unroll
        for (int kap = 0; kap < 2; ++kap)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float2 x = z[pad(m2 + M2 * (16 * kap + g + 4 * q))];
                d[kap * 4 + q] = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(x.x, x.y));
            }
        A[tp][0] = __builtin_bit_cast(half8, make_uint4(d[0], d[1], d[2], d[3]));
        A[tp][1] = __builtin_bit_cast(half8, make_uint4(d[4], d[5], d[6], d[7]));
    }
    f4 y[TPW][4];
#pragma unroll
    for (int tt = 0; tt < 4; ++tt) {
        const half8 bh0 = __builtin_bit_cast(half8, b1[(tt * 2 + 0) * 64 + l]);
        const half8 bl0 = __builtin_bit_cast(half8, b1[(tt * 2 + 1) * 64 + l]);
        const half8 bh1 = __builtin_bit_cast(half8, b1[((4 + tt) * 2 + 0) * 64 + l]);
        const half8 bl1 = __builtin_bit_cast(half8, b1[((4 + tt) * 2 + 1) * 64 + l]);
#pragma unroll
        for (int tp = 0; tp < TPW; ++tp) y[tp][tt] = mf_stage1(A[tp][0], A[tp][1], bh0, bl0, bh1, bl1);
    }
    __syncthreads();   // every wave's reads of z are done
#pragma unroll
    for (int tp = 0; tp < TPW; ++tp)
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m2 = 16 * (wv * TPW + tp) + 4 * g + i;
                z[pad(m2 * 32 + c + 16 * h)] = make_float2(y[tp][h][i], y[tp][2 + h][i]);
            }
    __syncthreads();
}

// Forward complex FFT of M = E*T points held as v[r] = z[t + T*r]; result
// in natural order in z.
template <int E, int T>
__device__ __forceinline__ void fft_forward(float2 *z, float2 *v, const float2 *tw, int t) {
    stockham_first<E, T>(z, v, t);
    stockham_rest<E, T, E>(z, tw, t);
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

// realfft's split method on an unpadded Z (natural order in global memory).
__device__ __forceinline__ float2 untangle_np(const float2 *zs, int k, int M, const float2 *tw_post) {
    if (k == 0 || k == M) {
        const float2 z0 = zs[0];
        return make_float2(k == 0 ? z0.x + z0.y : z0.x - z0.y, 0.0f);
    }
    const float2 A = zs[k], Bc = zs[M - k];
    const float ex = 0.5f * (A.x + Bc.x), ey = 0.5f * (A.y - Bc.y);
    const float dx = 0.5f * (A.x - Bc.x), dy = 0.5f * (A.y + Bc.y);
    const float2 w = tw_post[k];
    return make_float2(ex + (w.x * dy - w.y * dx), ey - (w.x * dx + w.y * dy));
}

// The inverse real FFT's input element k (< M) of the half spectrum X
// (M + 1 bins): Z'[k] = E' + i O' (the split method run backwards), returned
// conjugated for IFFT = conj(FFT(conj(.))).  DC/Nyquist imaginary parts are
// ignored (the reference zeroes them first).
__device__ __forceinline__ float2 irfft_pre_elem(const float2 *X, int k, int M,
                                                 const float2 *tw_post) {
    float2 z;
    if (k == 0) {
        z = make_float2(X[0].x + X[M].x, X[0].x - X[M].x);
    } else {
        const float2 A = X[k], Bc = X[M - k];
        const float er = A.x + Bc.x, ei = A.y - Bc.y;
        const float dr = A.x - Bc.x, di = A.y + Bc.y;
        const float2 w = tw_post[k];
        const float orr = dr * w.x - di * w.y;
        const float oi = dr * w.y + di * w.x;
        z = make_float2(er - oi, ei + orr);
    }
    return c_conj(z);
}

__device__ __forceinline__ float norm_sqr(float2 a) { return a.x * a.x + a.y * a.y; }

// Which half [0,T) / [T,2T) of the workgroup thread `tid` is in; made
// wave-uniform when a half spans whole waves (T >= 64).
template <int T>
__device__ __forceinline__ int half_index(int tid) {
    if constexpr (T >= 64) return __builtin_amdgcn_readfirstlane(tid / T);
    else return tid / T;
}

// ------------------------------------------------------- fused kernel
// Decoded, integer-shifted samples 2m and 2m+1 (one complex FFT input
// element) from the staged words: the two codes are adjacent in the bit
// stream, so one funnel shift of two words yields both.
__device__ __forceinline__ float2 decode_elem(const uint32_t *words, const StationDev &st,
                                              const float *lev_lds, int m, int shift, int N) {
    const int src0 = 2 * m - shift;
    const bool in0 = (unsigned)src0 < (unsigned)N;
    const bool in1 = (unsigned)(src0 + 1) < (unsigned)N;
    const int bp = src0 * st.bits;              // may be negative (arithmetic shift floors)
    const int w0 = bp >> 5;
    const uint32_t off = (uint32_t)bp & 31u;
    const uint32_t lo = (w0 >= 0 && w0 < st.nwords) ? words[w0] : 0u;
    const uint32_t hi = (w0 + 1 >= 0 && w0 + 1 < st.nwords) ? words[w0 + 1] : 0u;
    const uint32_t v = __builtin_amdgcn_alignbit(hi, lo, off);
    float x0 = level_of(v & st.code_mask, st, lev_lds);
    float x1 = level_of((v >> st.bits) & st.code_mask, st, lev_lds);
    if (st.lsb) {                               // LSB->USB flip of odd decode-order samples
        if (src0 & 1) x0 = -x0;
        else x1 = -x1;
    }
    return make_float2(in0 ? x0 : 0.0f, in1 ? x1 : 0.0f);
}


// E consecutive elements m0 .. m0 + E - 1 (samples 2 m0 .. 2 m0 + 2E - 1 of
// the shifted frame) of one station, with decode_elem's values: the run's
// words are read once and funnel-shifted into a bit stream whose codes sit
// at fixed offsets (BITS = 1, 2, 4, 8); samples outside the frame are zero,
// and with the LSB flip the samples of odd index in the unshifted frame are
// negated.  Per-sample range checks run only for runs that cross an edge.
template <int E, int BITS>
__device__ __forceinline__ void decode_run_t(const uint32_t *words, const StationDev &st,
                                             const float *lev_lds, int m0, int shift, int N,
                                             float2 (&v)[E]) {
    constexpr int NS = 2 * E, NW = (NS * BITS + 31) / 32;
    const int s0 = 2 * m0 - shift;                 // index of the run's first sample
    const int bp = s0 * BITS;                      // may be negative (arithmetic shift floors)
    const int w0 = bp >> 5;
    const uint32_t off = (uint32_t)bp & 31u;
    uint32_t w[NW + 1];
#pragma unroll
    for (int i = 0; i <= NW; ++i) {
        const int wi = w0 + i;
        w[i] = (wi >= 0 && wi < st.nwords) ? words[wi] : 0u;
    }
    uint32_t d[NW];
#pragma unroll
    for (int i = 0; i < NW; ++i) d[i] = __builtin_amdgcn_alignbit(w[i + 1], w[i], off);
    const bool inside = s0 >= 0 && s0 + NS <= N;
    const bool flip = st.lsb != 0;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        const uint32_t code = (d[(j * BITS) >> 5] >> ((j * BITS) & 31)) & st.code_mask;
        float x = level_of(code, st, lev_lds);
        if (flip && ((s0 + j) & 1)) x = -x;
        if (!inside && (unsigned)(s0 + j) >= (unsigned)N) x = 0.0f;
        if (j & 1) v[j >> 1].y = x;
        else v[j >> 1].x = x;
    }
}

template <int E>
__device__ __forceinline__ void decode_run(const uint32_t *words, const StationDev &st,
                                           const float *lev_lds, int m0, int shift, int N,
                                           float2 (&v)[E]) {
    switch (st.bits) {
    case 1: decode_run_t<E, 1>(words, st, lev_lds, m0, shift, N, v); break;
    case 2: decode_run_t<E, 2>(words, st, lev_lds, m0, shift, N, v); break;
    case 4: decode_run_t<E, 4>(words, st, lev_lds, m0, shift, N, v); break;
    case 8: decode_run_t<E, 8>(words, st, lev_lds, m0, shift, N, v); break;
    default:
#pragma unroll
        for (int r = 0; r < E; ++r) v[r] = decode_elem(words, st, lev_lds, m0 + r, shift, N);
    }
}

}  // namespace fxc
