// Chosen-form fused kernel for N = 1024 (M = 512 complex points per
// station-frame, z[m] = x[2m] + i x[2m+1]): a four-step FFT whose first
// stage runs on the matrix cores.
//
//   m = 16 m1 + m2,  k = k1 + 32 k2
//   stage 1 (FP16 MFMA, FP32 acc): Y[m2][k1]  = sum_m1 z[16 m1 + m2] W32^{m1 k1}
//   twiddle (FP32):                Y'[m2][k1] = Y[m2][k1] W512^{m2 k1}
//   stage 2 (FP32):                Z[k1 + 32 k2] = sum_m2 Y'[m2][k1] W16^{m2 k2}
//     k_sector_mf2: an FP32 MFMA GEMM; k_sector_mf3: one column per lane,
//     radix 4 x 4 on the VALU
//
// Stage 1 is the real GEMM D[m2][(k1,re|im)] = sum_(m1,re|im) A B with
// A = the samples (exact in FP16 for 1/2/4/8-bit levels) and B the DFT-32
// coefficients split as B = B_hi + B_lo (both FP16, B_lo = B - B_hi
// rounded, subnormal where small: B_hi + B_lo is within 2^-25 of B),
// both products accumulated in FP32.  Its accumulator layout (lane l: rows m2 = 4(l/16)+r,
// column l%16) is exactly the B-operand layout stage 2 needs (K = (m2,
// re|im), column k1), so the data never leaves registers between stages.
//
// One wave carries a whole frame pair (both stations): decode into a
// zero-padded FP16 staging row (the integer shift and its zero fill become
// an offset into that row), both stages for both stations, the untangle to
// the real spectrum in bin pairs (k, M-k), and the ACF/XCF products into
// FP64 registers.  Eight waves per workgroup share the read-only tables and
// never wait for each other inside the frame loop.
#include "fxc_dev.h"

namespace fxc {

namespace mf {
constexpr int N = 1024, M = 512;
constexpr int WAVES = 8;                 // waves per workgroup
constexpr int PAD = 64;                  // zero FP16 samples each side of a staged frame
constexpr int SROW = N + 2 * PAD;        // halves per staged station-frame
constexpr int FB = 64;                   // frames whose parameters one batch computes
}  // namespace mf

struct MfShared {
    uint4 b1[16 * 64];                           // stage-1 B operands [idx][lane]
    float a2[16 * 64];                           // stage-2 A operands [rt*8 + b][lane]
    float tw2[16 * 64];                          // W512^{m2 k1} [(ct*4 + i)*2 + part][lane]
    float twu[8 * 64];                           // W1024^p [2j + part][lane]
    double mid[mf::WAVES][4];                    // per-wave accumulators of bin M/2
    uint32_t slut[2][4 * 256];                   // generic shuffle: byte -> output bits
    uint16_t dlut[2][MFT_DLUT_HALVES];           // byte -> FP16 samples
    uint16_t stage[mf::WAVES][2][mf::SROW];      // [wave][station] decoded frame + zero pad
    float2 z[mf::WAVES][2][mf::M];               // [wave][station] Z/2 in natural order
};

// Orders this wave's LDS writes before its later LDS reads (a wave's LDS
// operations complete in issue order; this keeps the compiler from moving
// them across each other).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, lane);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), lane);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ float readlane_f(float v, int lane) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(uint32_t, v), lane));
}

// (cos, sin)(2 pi r) for |r| <= 0.5 turns: quadrant reduction to
// |t| <= 1/8 turn and Taylor polynomials through x^9 / x^8 (truncation
// error below 3e-8 on |x| <= pi/4).
__device__ __forceinline__ float2 sincos_turns(float r) {
    const float q = rintf(4.0f * r);
    const float t = r - 0.25f * q;
    const float x = t * 6.28318530717958647692f;
    const float x2 = x * x;
    float s = fmaf(x2, 2.75573192e-6f, -1.98412698e-4f);
    s = fmaf(x2, s, 8.33333333e-3f);
    s = fmaf(x2, s, -1.66666667e-1f);
    s = fmaf(x2 * x, s, x);
    float c = fmaf(x2, 2.48015873e-5f, -1.38888889e-3f);
    c = fmaf(x2, c, 4.16666667e-2f);
    c = fmaf(x2, c, -0.5f);
    c = fmaf(x2, c, 1.0f);
    const int qi = ((int)q) & 3;
    float co = c, si = s;
    if (qi == 1) { co = -s; si = c; }
    else if (qi == 2) { co = -c; si = -s; }
    else if (qi == 3) { co = s; si = -c; }
    return make_float2(co, si);
}


// Per-frame values of the chosen form (main.rs:184-241, 133-149 restated in
// turns): integer shifts, the fringe difference phw = w1 - w2 and the
// per-bin XCF phase step dth = -df (frac1 - frac2), so that the XCF phasor
// of grid bin k is exp(2 pi i (phw + k dth)) when both stations share the
// band start (a1s == a2s).  F1 = exp(2 pi i 64 dth), C = exp(2 pi i (2 phw
// + M dth)) turn the per-lane phasor into the phasors of its other bins.
struct MfFrame {
    int32_t s1, s2;
    double phw, dth;
    float2 F1, C;
};

__device__ MfFrame mf_frame(int64_t idx, const SectorArgs &a) {
    MfFrame o;
    double tau1, tau2, tf1, tf2;
    frame_taus(idx, a.dly, tau1, tau2, tf1, tf2);
    int64_t n1, n2;
    double fr1, fr2;
    split_delay(tau1, a.dly.fs, n1, fr1);
    split_delay(tau2, a.dly.fs, n2, fr2);
    o.s1 = clamp_shift(n1, a.N);
    o.s2 = clamp_shift(n2, a.N);
    const double w1 = a.dly.lo1 == 0.0 ? 0.0 : wrap_turns(-a.dly.lo1 * tf1);
    const double w2 = a.dly.lo2 == 0.0 ? 0.0 : wrap_turns(-a.dly.lo2 * tf2);
    o.phw = w1 - w2;
    o.dth = -a.df * (fr1 - fr2);
    o.F1 = sincos_turns((float)wrap_turns(64.0 * o.dth));
    o.C = sincos_turns((float)wrap_turns(2.0 * o.phw + (double)a.M * o.dth));
    return o;
}

// ------------------------------------------------------------ decode
__device__ __forceinline__ uint32_t mf_shuffle(uint32_t v, int kind, const uint32_t *slut) {
    if (kind == SHUF_PAIRSWAP) return ((v & 0xAAAAAAAAu) >> 1) | ((v & 0x55555555u) << 1);
    if (kind == SHUF_GENERIC)
        return slut[v & 255u] | slut[256 + ((v >> 8) & 255u)] | slut[512 + ((v >> 16) & 255u)] |
               slut[768 + (v >> 24)];
    return v;
}

// The packed words of one station-frame this lane stages (word l + 64 u).
__device__ __forceinline__ void mf_load(const uint8_t *frame, int nwords, int l, uint32_t w[4]) {
    const uint32_t *src = reinterpret_cast<const uint32_t *>(frame);
#pragma unroll
    for (int u = 0; u < 4; ++u)
        if (l + 64 * u < nwords) w[u] = __builtin_nontemporal_load(src + l + 64 * u);
}

// One packed word w (word wi of its frame) decoded into FP16 samples at their
// frame positions (utils.rs:215-376: LSB-first codes, levels, odd samples
// negated for LSB stations -- the sign is folded into the per-station byte
// table).
__device__ __forceinline__ void mf_stage_word(uint32_t w, int bits, int kind, const uint32_t *slut,
                                              const uint16_t *dlut, uint16_t *frame, int wi) {
    {
        const uint32_t v = mf_shuffle(w, kind, slut);
        const uint32_t b0 = v & 255u, b1 = (v >> 8) & 255u, b2 = (v >> 16) & 255u, b3 = v >> 24;
        if (bits == 2) {
            const uint2 e0 = *reinterpret_cast<const uint2 *>(dlut + 4 * b0);
            const uint2 e1 = *reinterpret_cast<const uint2 *>(dlut + 4 * b1);
            const uint2 e2 = *reinterpret_cast<const uint2 *>(dlut + 4 * b2);
            const uint2 e3 = *reinterpret_cast<const uint2 *>(dlut + 4 * b3);
            uint4 *dst = reinterpret_cast<uint4 *>(frame + 16 * wi);
            dst[0] = make_uint4(e0.x, e0.y, e1.x, e1.y);
            dst[1] = make_uint4(e2.x, e2.y, e3.x, e3.y);
        } else if (bits == 1) {
            uint4 *dst = reinterpret_cast<uint4 *>(frame + 32 * wi);
            dst[0] = *reinterpret_cast<const uint4 *>(dlut + 8 * b0);
            dst[1] = *reinterpret_cast<const uint4 *>(dlut + 8 * b1);
            dst[2] = *reinterpret_cast<const uint4 *>(dlut + 8 * b2);
            dst[3] = *reinterpret_cast<const uint4 *>(dlut + 8 * b3);
        } else if (bits == 4) {
            const uint32_t *d32 = reinterpret_cast<const uint32_t *>(dlut);
            *reinterpret_cast<uint4 *>(frame + 8 * wi) = make_uint4(d32[b0], d32[b1], d32[b2], d32[b3]);
        } else {  // 8 bits: one sample per byte; odd bytes are odd samples
            const uint32_t h0 = dlut[b0], h1 = dlut[256 + b1], h2 = dlut[b2], h3 = dlut[256 + b3];
            *reinterpret_cast<uint2 *>(frame + 4 * wi) = make_uint2(h0 | (h1 << 16), h2 | (h3 << 16));
        }
    }
}

// This lane's words (l + 64 u) of one station-frame.
__device__ __forceinline__ void mf_stage(const uint32_t w[4], int bits, int kind, int nwords,
                                         const uint32_t *slut, const uint16_t *dlut,
                                         uint16_t *frame, int l) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int wi = l + 64 * u;
        if (wi < nwords) mf_stage_word(w[u], bits, kind, slut, dlut, frame, wi);
    }
}

// Stage-1 A operands of one station-frame.  Lane l (m2 = l%16, g = l/16)
// holds, in K-step kap, slots s = 2j + e: sample 2(16 m1 + m2) + e of the
// shifted frame with m1 = 16 kap + g + 4j, i.e. the FP16 pair at staged
// position PAD + 32 m1 + 2 m2 - sh.  (The lane groups g of a half-wave then
// read 16 dwords apart: no LDS bank conflict.)  Positions outside the
// staged row read the zero padding.
// R2 rows m2 per station-frame (M = 32 R2, m = R2 m1 + m2); row tile rt
// covers m2 = 16 rt + (l % 16).  SROW: halves per staged row.
template <int R2, int SROW>
__device__ __forceinline__ void mf_build_a_t(const uint16_t *row, int sh, int l, int rt, half8 &a0,
                                             half8 &a1) {
    constexpr int OK = 16 * R2, OJ = 4 * R2;   // dword offsets of K-step kap and slot pair j
    const uint32_t *r32 = reinterpret_cast<const uint32_t *>(row);
    const int base = mf::PAD + 2 * ((l & 15) + 16 * rt) - sh + 2 * R2 * (l >> 4);
    uint32_t d[8];
    const bool near = (sh >= -mf::PAD) && (sh <= mf::PAD);
    if (near && (sh & 1) == 0) {
        // base is even: one dword pointer, constant offsets
        const uint32_t *q = r32 + (base >> 1);
#pragma unroll
        for (int k = 0; k < 8; ++k) d[k] = q[OK * (k >> 2) + OJ * (k & 3)];
    } else if (near) {
        // odd shift: the FP16 pair straddles two dwords
        const uint32_t *q = r32 + ((base - 1) >> 1);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int o = OK * (k >> 2) + OJ * (k & 3);
            d[k] = __builtin_amdgcn_alignbyte(q[o + 1], q[o], 2);
        }
    } else if ((sh & 1) == 0) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            int p = base + 2 * OK * (k >> 2) + 2 * OJ * (k & 3);
            p = (p >= 0 && p < SROW - 1) ? p : 0;
            d[k] = r32[p >> 1];
        }
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            int p = base + 2 * OK * (k >> 2) + 2 * OJ * (k & 3);
            p = (p >= 0 && p < SROW - 1) ? p : 1;
            d[k] = __builtin_amdgcn_alignbyte(r32[(p + 1) >> 1], r32[(p - 1) >> 1], 2);
        }
    }
    a0 = __builtin_bit_cast(half8, make_uint4(d[0], d[1], d[2], d[3]));
    a1 = __builtin_bit_cast(half8, make_uint4(d[4], d[5], d[6], d[7]));
}

__device__ __forceinline__ void mf_build_a(const uint16_t *row, int sh, int l, half8 &a0, half8 &a1) {
    mf_build_a_t<16, mf::SROW>(row, sh, l, 0, a0, a1);
}

// ------------------------------------------------------------- FFT
__device__ __forceinline__ half8 ld_b1(const uint4 *b1, int idx, int l) {
    return __builtin_bit_cast(half8, b1[idx * 64 + l]);
}

// Twiddle and stage 2 of one station-frame; writes Z/2 in natural order.
// y[t][r] from stage 1; tw2 / a2 are the [idx][lane] tables.
__device__ __forceinline__ void mf_stage2(const f4 y[4], const float *tw2, const float *a2,
                                          float2 *z, int l) {
    float2 yc[2][4];
#pragma unroll
    for (int ct = 0; ct < 2; ++ct)
#pragma unroll
        for (int i = 0; i < 4; ++i)
            yc[ct][i] = c_mul(make_float2(y[ct][i], y[2 + ct][i]),
                              make_float2(tw2[((ct * 4 + i) * 2) * 64 + l],
                                          tw2[((ct * 4 + i) * 2 + 1) * 64 + l]));
    f4 d[2][2];
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) d[rt][ct] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int b = 0; b < 8; ++b) {
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
            const float bv = (b & 1) ? yc[ct][b >> 1].y : yc[ct][b >> 1].x;
#pragma unroll
            for (int rt = 0; rt < 2; ++rt)
                d[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a2[(rt * 8 + b) * 64 + l], bv,
                                                                 d[rt][ct], 0, 0, 0);
        }
    }
#pragma unroll
    for (int ct = 0; ct < 2; ++ct)
#pragma unroll
        for (int r = 0; r < 4; ++r)
            z[(l & 15) + 16 * ct + 32 * (4 * (l >> 4) + r)] = make_float2(d[0][ct][r], d[1][ct][r]);
}

// Stage 2 through a transposed copy in the Z slot (producer/consumer
// organization).  Each station's twiddled stage-1 output Y'[m2][k1] is
// stored column by column (the 16 values m2 = 0..15 of column k1 in a row
// of 128 bytes, 16-byte pairs XOR-swizzled by k1 & 7, so that any 8
// consecutive columns -- one pass of a 16-byte store or load over the 32
// LDS banks -- fall on 8 different 16-byte bank groups); then lane l takes
// column k1 = l & 31 of station l >> 5 and runs the whole DFT-16 in its own
// registers (radix 4 x 4, no cross-lane traffic), writing Z/2 over the same
// slot in natural order.
__device__ __forceinline__ int mf_tidx(int s, int k1, int pair) {
    return s * mf::M + k1 * 16 + 2 * (pair ^ (k1 & 7));
}

// W512^{m2 k1} of this lane's rows m2 = 4g + r in column half h.
__device__ __forceinline__ void mf_load_tw(const float *tw2, int h, int l, float2 tw[4]) {
#pragma unroll
    for (int r = 0; r < 4; ++r)
        tw[r] = make_float2(tw2[((h * 4 + r) * 2) * 64 + l], tw2[((h * 4 + r) * 2 + 1) * 64 + l]);
}

// Twiddle of column half h (k1 = l%16 + 16h; re / im from stage-1 tiles h
// and 2 + h) of station s, stored transposed.
__device__ __forceinline__ void mf_twiddle_store_half(f4 re, f4 im, const float2 tw[4], float2 *slot,
                                                      int s, int h, int l) {
    float2 v[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = c_mul(make_float2(re[r], im[r]), tw[r]);
    const int g = l >> 4, k1 = (l & 15) + 16 * h;   // rows m2 = 4g + r: pairs 2g and 2g + 1
    *reinterpret_cast<float4 *>(slot + mf_tidx(s, k1, 2 * g)) = make_float4(v[0].x, v[0].y, v[1].x, v[1].y);
    *reinterpret_cast<float4 *>(slot + mf_tidx(s, k1, 2 * g + 1)) = make_float4(v[2].x, v[2].y, v[3].x, v[3].y);
}

__device__ __forceinline__ void mf_twiddle_store(const f4 y[4], const float *tw2, float2 *slot, int s,
                                                 int l) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        float2 tw[4];
        mf_load_tw(tw2, h, l, tw);
        mf_twiddle_store_half(y[h], y[2 + h], tw, slot, s, h, l);
    }
}

// x * W16^n for the constant twiddles of the DFT-16 (W16 = exp(-2 pi i / 16)).
template <int n>
__device__ __forceinline__ float2 mf_w16(float2 x) {
    constexpr float c1 = 0.92387953251128674f, s1 = 0.38268343236508978f, c2 = 0.70710678118654752f;
    if constexpr (n == 4) {
        return make_float2(x.y, -x.x);
    } else {
        constexpr float wr = n == 1 ? c1 : n == 2 ? c2 : n == 3 ? s1 : n == 6 ? -c2 : -c1;
        constexpr float wi = n == 1 ? -s1 : n == 2 ? -c2 : n == 3 ? -c1 : n == 6 ? -c2 : s1;
        return make_float2(x.x * wr - x.y * wi, x.x * wi + x.y * wr);
    }
}

// DFT-4 of (x0, x1, x2, x3) with W4 = -i, outputs in natural order.
__device__ __forceinline__ void mf_dft4(float2 x0, float2 x1, float2 x2, float2 x3, float2 &o0,
                                        float2 &o1, float2 &o2, float2 &o3) {
    const float2 a0 = make_float2(x0.x + x2.x, x0.y + x2.y), a1 = make_float2(x0.x - x2.x, x0.y - x2.y);
    const float2 b0 = make_float2(x1.x + x3.x, x1.y + x3.y), b1 = make_float2(x1.x - x3.x, x1.y - x3.y);
    o0 = make_float2(a0.x + b0.x, a0.y + b0.y);
    o2 = make_float2(a0.x - b0.x, a0.y - b0.y);
    o1 = make_float2(a1.x + b1.y, a1.y - b1.x);   // a1 - i b1
    o3 = make_float2(a1.x - b1.y, a1.y + b1.x);   // a1 + i b1
}

// DFT-16 of x[0..15] in registers (radix 4 x 4: input n = 4g + r, output
// k = a + 4b, X[a + 4b] = sum_r W4^{rb} W16^{ra} sum_g W4^{ga} x[4g + r]);
// X[k] is written to zo[stride * k].
__device__ __forceinline__ void mf_dft16_out(const float2 (&x)[16], float2 *zo, int stride) {
    float2 u[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r) mf_dft4(x[r], x[4 + r], x[8 + r], x[12 + r], u[0][r], u[1][r], u[2][r], u[3][r]);
    u[1][1] = mf_w16<1>(u[1][1]);
    u[1][2] = mf_w16<2>(u[1][2]);
    u[1][3] = mf_w16<3>(u[1][3]);
    u[2][1] = mf_w16<2>(u[2][1]);
    u[2][2] = mf_w16<4>(u[2][2]);
    u[2][3] = mf_w16<6>(u[2][3]);
    u[3][1] = mf_w16<3>(u[3][1]);
    u[3][2] = mf_w16<6>(u[3][2]);
    u[3][3] = mf_w16<9>(u[3][3]);
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        float2 o0, o1, o2, o3;
        mf_dft4(u[a][0], u[a][1], u[a][2], u[a][3], o0, o1, o2, o3);
        zo[stride * a] = o0;
        zo[stride * (a + 4)] = o1;
        zo[stride * (a + 8)] = o2;
        zo[stride * (a + 12)] = o3;
    }
}

// Column DFT-16 of both stations: reads the transposed Y', writes Z/2.
__device__ __forceinline__ void mf_columns_dft16(float2 *slot, int l) {
    const int s = l >> 5, k1 = l & 31;
    float2 x[16];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const float4 q = *reinterpret_cast<const float4 *>(slot + mf_tidx(s, k1, j));
        x[2 * j] = make_float2(q.x, q.y);
        x[2 * j + 1] = make_float2(q.z, q.w);
    }
    wave_lds_sync();   // every lane's column is read before Z overwrites the slot
    mf_dft16_out(x, slot + s * mf::M + k1, 32);
}

// Real-spectrum bins p and M-p from Z/2 (realfft's split method, paired):
// E = Z'[p] + conj Z'[M-p], O = Z'[p] - conj Z'[M-p], T = W_N^p O,
// X[p] = E - iT, X[M-p] = conj(E + iT).  p = 0 gives DC and Nyquist.
__device__ __forceinline__ void untangle_pair(float2 a, float2 b, float2 w, float2 &xp, float2 &xm) {
    const float ex = a.x + b.x, ey = a.y - b.y;
    const float ox = a.x - b.x, oy = a.y + b.y;
    const float tx = w.x * ox - w.y * oy, ty = w.x * oy + w.y * ox;
    xp = make_float2(ex + ty, ey - tx);
    xm = make_float2(ex - ty, -(ey + tx));
}

// ------------------------------------------------------- fused kernel
// Bins of lane l: slot 2j -> p = l + 64 j, slot 2j+1 -> M - p (j < 4);
// slot 8 -> M/2 (lane 0 only).
__device__ __forceinline__ int mf_bin(int slot, int l) {
    if (slot == 8) return mf::M / 2;
    const int p = l + 64 * (slot >> 1);
    return (slot & 1) ? mf::M - p : p;
}

__global__ __launch_bounds__(512) void k_sector_mf2(SectorArgs a) {
    __shared__ MfShared sh;
    constexpr int M = mf::M;
    const int tid = threadIdx.x;
    const int l = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint8_t *tab = reinterpret_cast<const uint8_t *>(a.mf_tab);

    // ---- read-only tables and the zero padding
    {
        const uint4 *g = reinterpret_cast<const uint4 *>(tab + MFT_B1);
        for (int i = tid; i < 16 * 64; i += 512) sh.b1[i] = g[i];
        const uint32_t *gs = reinterpret_cast<const uint32_t *>(tab + MFT_SLUT);
        for (int i = tid; i < 2 * 4 * 256; i += 512) (&sh.slut[0][0])[i] = gs[i];
        const uint32_t *gd = reinterpret_cast<const uint32_t *>(tab + MFT_DLUT);
        for (int i = tid; i < 2 * MFT_DLUT_HALVES / 2; i += 512)
            reinterpret_cast<uint32_t *>(&sh.dlut[0][0])[i] = gd[i];
        uint32_t *st32 = reinterpret_cast<uint32_t *>(&sh.stage[0][0][0]);
        for (int i = tid; i < mf::WAVES * 2 * mf::SROW / 2; i += 512) st32[i] = 0u;
        const float *ga = reinterpret_cast<const float *>(tab + MFT_A2);
        const float *gt = reinterpret_cast<const float *>(tab + MFT_TW2);
        const float *gu = reinterpret_cast<const float *>(tab + MFT_TWU);
        for (int i = tid; i < 16 * 64; i += 512) {
            sh.a2[i] = ga[i];
            sh.tw2[i] = gt[i];
        }
        for (int i = tid; i < 8 * 64; i += 512) sh.twu[i] = gu[i];
        if (tid < mf::WAVES * 4) (&sh.mid[0][0])[tid] = 0.0;
    }
    __syncthreads();

    const bool do_acf = (a.flags & FLAG_ACF) != 0;
    const bool do_xcf = (a.flags & FLAG_XCF) != 0;
    const bool ov_only = (a.flags & FLAG_ACF_OVERLAP) != 0;
    const int gs = a.a1s, ge = a.a1s + a.ov;   // a1s == a2s on this path
    // per-slot masks: grid-station ACF, other-station ACF / XCF (overlap)
    uint32_t m_grid = 0, m_ov = 0;
#pragma unroll
    for (int s = 0; s < 9; ++s) {
        const int k = mf_bin(s, l);
        const bool in_ov = k >= gs && k < ge;
        if (!ov_only || in_ov) m_grid |= 1u << s;
        if (in_ov) m_ov |= 1u << s;
    }
    const uint32_t m11 = a.grid == 0 ? m_grid : m_ov;
    const uint32_t m22 = a.grid == 0 ? m_ov : m_grid;

    double acc[8][4];
#pragma unroll
    for (int s = 0; s < 8; ++s)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[s][c] = 0.0;

    const int bits1 = a.st[0].bits, bits2 = a.st[1].bits;
    const int kind1 = a.st[0].shuffle_kind, kind2 = a.st[1].shuffle_kind;
    const int nw1 = a.st[0].nwords, nw2 = a.st[1].nwords;
    const int64_t bpf1 = a.st[0].bpf, bpf2 = a.st[1].bpf;
    uint16_t *row1 = sh.stage[wv][0];
    uint16_t *row2 = sh.stage[wv][1];
    float2 *z1 = sh.z[wv][0];
    float2 *z2 = sh.z[wv][1];

    // frames of this wave: a contiguous slice of the workgroup's range
    const int64_t wg0 = (int64_t)blockIdx.x * a.frames_per_wg;
    int64_t wg1 = wg0 + a.frames_per_wg;
    if (wg1 > a.nf) wg1 = a.nf;
    const int64_t per_wave = (a.frames_per_wg + mf::WAVES - 1) / mf::WAVES;
    const int64_t f0 = wg0 + per_wave * wv;
    const int64_t f1 = (f0 + per_wave < wg1) ? f0 + per_wave : wg1;

    // Software pipeline: the staging rows hold frame f while its FFTs run;
    // frame f+1 is decoded into them during frame f's products (disjoint LDS
    // regions), and the words of frame f+2 are in flight from HBM.
    uint32_t cur1[4] = {0, 0, 0, 0}, cur2[4] = {0, 0, 0, 0};   // words of the next frame
    if (f0 < f1) {
        mf_load(a.raw1 + f0 * bpf1, nw1, l, cur1);
        mf_load(a.raw2 + f0 * bpf2, nw2, l, cur2);
        mf_stage(cur1, bits1, kind1, nw1, sh.slut[0], sh.dlut[0], row1 + mf::PAD, l);
        mf_stage(cur2, bits2, kind2, nw2, sh.slut[1], sh.dlut[1], row2 + mf::PAD, l);
        if (f0 + 1 < f1) {
            mf_load(a.raw1 + (f0 + 1) * bpf1, nw1, l, cur1);
            mf_load(a.raw2 + (f0 + 1) * bpf2, nw2, l, cur2);
        }
        wave_lds_sync();
    }
    for (int64_t fb = f0; fb < f1; fb += mf::FB) {
        const int nb = (int)((f1 - fb) < mf::FB ? (f1 - fb) : mf::FB);
        MfFrame pf = {};
        if (l < nb) pf = mf_frame(a.first_frame + fb + l, a);
        for (int jf = 0; jf < nb; ++jf) {
            const int64_t f = fb + jf;
            uint32_t nxt1[4] = {0, 0, 0, 0}, nxt2[4] = {0, 0, 0, 0};
            if (f + 2 < f1) {
                mf_load(a.raw1 + (f + 2) * bpf1, nw1, l, nxt1);
                mf_load(a.raw2 + (f + 2) * bpf2, nw2, l, nxt2);
            }
            const int s1 = __builtin_amdgcn_readlane(pf.s1, jf);
            const int s2 = __builtin_amdgcn_readlane(pf.s2, jf);

            half8 a10, a11, a20, a21;
            mf_build_a(row1, s1, l, a10, a11);
            mf_build_a(row2, s2, l, a20, a21);

            // ---- stage 1 (B operands shared by both stations)
            f4 y1[4], y2[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const half8 bh0 = ld_b1(sh.b1, (0 * 4 + t) * 2 + 0, l);
                const half8 bl0 = ld_b1(sh.b1, (0 * 4 + t) * 2 + 1, l);
                const half8 bh1 = ld_b1(sh.b1, (1 * 4 + t) * 2 + 0, l);
                const half8 bl1 = ld_b1(sh.b1, (1 * 4 + t) * 2 + 1, l);
                y1[t] = mf_stage1(a10, a11, bh0, bl0, bh1, bl1);
                y2[t] = mf_stage1(a20, a21, bh0, bl0, bh1, bl1);
            }
            // ---- twiddle + stage 2, Z/2 to LDS
            mf_stage2(y1, sh.tw2, sh.a2, z1, l);
            mf_stage2(y2, sh.tw2, sh.a2, z2, l);
            wave_lds_sync();   // Z visible; this frame's staging rows no longer read

            // ---- decode frame f+1 into the staging rows (overlaps the products)
            if (f + 1 < f1) {
                mf_stage(cur1, bits1, kind1, nw1, sh.slut[0], sh.dlut[0], row1 + mf::PAD, l);
                mf_stage(cur2, bits2, kind2, nw2, sh.slut[1], sh.dlut[1], row2 + mf::PAD, l);
            }

            // ---- XCF phasor of this lane's first bin: FP64 angle reduced mod
            // 1 turn, then the hardware sin/cos of 2 pi r (|r| <= 1/2 turn)
            const double phw = readlane_d(pf.phw, jf);
            const double dth = readlane_d(pf.dth, jf);
            const float2 F1 = make_float2(readlane_f(pf.F1.x, jf), readlane_f(pf.F1.y, jf));
            const float2 Cm = make_float2(readlane_f(pf.C.x, jf), readlane_f(pf.C.y, jf));
            const float r0 = (float)wrap_turns(fma((double)l, dth, phw));
            float2 P = make_float2(__builtin_amdgcn_cosf(r0), __builtin_amdgcn_sinf(r0));

            // ---- products on the output grid (main.rs:2491-2568)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int p = l + 64 * j;
                const int pm = (M - p) & (M - 1);
                const float2 w = make_float2(sh.twu[(2 * j) * 64 + l], sh.twu[(2 * j + 1) * 64 + l]);
                float2 x1p, x1m, x2p, x2m;
                untangle_pair(z1[p], z1[pm], w, x1p, x1m);
                untangle_pair(z2[p], z2[pm], w, x2p, x2m);
                const int sp = 2 * j, sm = 2 * j + 1;
                if (do_acf) {
                    acc[sp][0] += (double)(((m11 >> sp) & 1u) ? norm_sqr(x1p) : 0.0f);
                    acc[sp][1] += (double)(((m22 >> sp) & 1u) ? norm_sqr(x2p) : 0.0f);
                    acc[sm][0] += (double)(((m11 >> sm) & 1u) ? norm_sqr(x1m) : 0.0f);
                    acc[sm][1] += (double)(((m22 >> sm) & 1u) ? norm_sqr(x2m) : 0.0f);
                }
                if (do_xcf) {
                    const float2 Pp = ((m_ov >> sp) & 1u) ? P : make_float2(0.f, 0.f);
                    const float2 Pm = ((m_ov >> sm) & 1u) ? c_mul(Cm, c_conj(P)) : make_float2(0.f, 0.f);
                    const float2 vp = c_mul(c_mul(x1p, c_conj(x2p)), Pp);
                    const float2 vm = c_mul(c_mul(x1m, c_conj(x2m)), Pm);
                    acc[sp][2] += (double)vp.x;
                    acc[sp][3] += (double)vp.y;
                    acc[sm][2] += (double)vm.x;
                    acc[sm][3] += (double)vm.y;
                }
                P = c_mul(P, F1);
            }
            if (l == 0) {   // k = M/2 pairs with itself: X = 2 conj(Z'[M/2]); P is P(M/2)
                const float2 u1 = z1[M / 2], u2 = z2[M / 2];
                const float2 x1 = make_float2(2.0f * u1.x, -2.0f * u1.y);
                const float2 x2 = make_float2(2.0f * u2.x, -2.0f * u2.y);
                double *mid = sh.mid[wv];
                if (do_acf) {
                    mid[0] += (double)(((m11 >> 8) & 1u) ? norm_sqr(x1) : 0.0f);
                    mid[1] += (double)(((m22 >> 8) & 1u) ? norm_sqr(x2) : 0.0f);
                }
                if (do_xcf) {
                    const float2 Pp = ((m_ov >> 8) & 1u) ? P : make_float2(0.f, 0.f);
                    const float2 v = c_mul(c_mul(x1, c_conj(x2)), Pp);
                    mid[2] += (double)v.x;
                    mid[3] += (double)v.y;
                }
            }
            wave_lds_sync();
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                cur1[u] = nxt1[u];
                cur2[u] = nxt2[u];
            }
        }
    }

    // ---- workgroup partial: waves add their bins in a fixed order
    __syncthreads();
    double *red = reinterpret_cast<double *>(&sh.z[0][0][0]);   // [4][M + 1]
    constexpr int H = M + 1;
    for (int w = 0; w < mf::WAVES; ++w) {
        if (wv == w) {
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const int k = mf_bin(s, l);
#pragma unroll
                for (int c = 0; c < 4; ++c) red[c * H + k] = (w == 0 ? 0.0 : red[c * H + k]) + acc[s][c];
            }
            if (l == 0)
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    red[c * H + M / 2] = (w == 0 ? 0.0 : red[c * H + M / 2]) + sh.mid[w][c];
        }
        __syncthreads();
    }
    double *out = a.part + (size_t)blockIdx.x * kNumProducts * H;
    for (int k = tid; k < H; k += 512) {
        out[P_PH * H + k] = 0.0;
        out[P_11 * H + k] = red[0 * H + k];
        out[P_22 * H + k] = red[1 * H + k];
        out[P_12R * H + k] = red[2 * H + k];
        out[P_12I * H + k] = red[3 * H + k];
    }
}

// ----------------------------------------- producer/consumer variant
// The same per-frame work split by role over 12 waves (768 threads): waves
// 0..7 are producers, 8..11 consumers.  Wave w runs on SIMD w % 4, so each
// SIMD holds producers c and c + 4 and their consumer 8 + c.  A producer
// decodes its frames, builds the A operands and runs stage 1 on the matrix
// cores for both stations, storing the twiddled result transposed in its
// LDS slot; then each lane runs one column's DFT-16 on the VALU
// (mf_columns_dft16; on gfx950 the FP32 MFMA and the VALU do not run
// concurrently, and the radix-4 x 4 DFT-16 takes fewer cycles than the
// dense FP32 GEMM) and writes Z/2 over the slot; the consumer untangles and accumulates the products of both its producers'
// frames, alternating between them.  Two producers per SIMD hide each
// other's matrix-core latency and the consumer's VALU work fills the matrix
// shadow; no wave carries the other role's registers.
//
// Hand-over per producer p (one slot): prod[p] counts frames published,
// cons[p] frames whose slot its consumer has released (workgroup-scope
// release / acquire on LDS).  Each wait is bounded (kMfSpinLimit polls); a
// wait cannot fail while the waves run, the bound only guarantees that
// every wave finishes.  A wave whose wait runs out stops waiting for the
// rest of its frames and marks the workgroup broken, whose partials are
// then written as NaN so the failure cannot pass as a result.
namespace mf {
constexpr int PROD = 8, CONS = 4;
constexpr int THREADS3 = 64 * (PROD + CONS);
// staging row of the producer/consumer kernels: one N = 1024 station-frame
// with its padding (SROW halves), or four N = 256 frames with theirs
constexpr int SROW3 = 1280;
}  // namespace mf
constexpr uint32_t kMfSpinLimit = 1u << 22;

struct MfShared3 {
    uint4 b1[16 * 64];
    float a2[16 * 64];
    float tw2[16 * 64];
    float twu[8 * 64];
    uint32_t slut[2][4 * 256];
    uint16_t dlut[2][MFT_DLUT_HALVES];
    uint16_t stage[mf::PROD][2][mf::SROW3];    // producers' staging rows
    alignas(16) float2 z[mf::PROD][2][mf::M];  // Z/2 slot per producer (16-byte stores)
    int32_t prod[mf::PROD], cons[mf::PROD];
    int32_t broken;
};

// Waits until *flag >= v; false if the bound ran out (or an earlier wait
// of this wave did: `ok` is cleared and later waits return at once).
__device__ __forceinline__ bool mf_wait_ge(int32_t *flag, int32_t v, bool &ok, int32_t *broken) {
    if (!ok) return false;
    for (uint32_t spin = 0; spin < kMfSpinLimit; ++spin) {
        const int32_t c = __builtin_amdgcn_readfirstlane(
            __hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
        if (c >= v) return true;
        __builtin_amdgcn_s_sleep(3);
    }
    ok = false;
    __hip_atomic_store(broken, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    return false;
}

__device__ __forceinline__ void mf_publish(int32_t *flag, int32_t v) {
    __hip_atomic_store(flag, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ void mf3_producer(const SectorArgs &a, MfShared3 &sh, int p, int l, int64_t f0,
                             int64_t f1) {
    const int bits1 = a.st[0].bits, bits2 = a.st[1].bits;
    const int kind1 = a.st[0].shuffle_kind, kind2 = a.st[1].shuffle_kind;
    const int nw1 = a.st[0].nwords, nw2 = a.st[1].nwords;
    const int64_t bpf1 = a.st[0].bpf, bpf2 = a.st[1].bpf;
    uint16_t *row1 = sh.stage[p][0];
    uint16_t *row2 = sh.stage[p][1];
    uint32_t cur1[4] = {0, 0, 0, 0}, cur2[4] = {0, 0, 0, 0};
    bool ok = true;
    if (f0 < f1) {
        mf_load(a.raw1 + f0 * bpf1, nw1, l, cur1);
        mf_load(a.raw2 + f0 * bpf2, nw2, l, cur2);
        mf_stage(cur1, bits1, kind1, nw1, sh.slut[0], sh.dlut[0], row1 + mf::PAD, l);
        mf_stage(cur2, bits2, kind2, nw2, sh.slut[1], sh.dlut[1], row2 + mf::PAD, l);
        if (f0 + 1 < f1) {
            mf_load(a.raw1 + (f0 + 1) * bpf1, nw1, l, cur1);
            mf_load(a.raw2 + (f0 + 1) * bpf2, nw2, l, cur2);
        }
        wave_lds_sync();
    }
    for (int64_t fb = f0; fb < f1; fb += mf::FB) {
        const int nb = (int)((f1 - fb) < mf::FB ? (f1 - fb) : mf::FB);
        MfFrame pf = {};
        if (l < nb) pf = mf_frame(a.first_frame + fb + l, a);
        for (int jf = 0; jf < nb; ++jf) {
            const int64_t f = fb + jf;
            const int32_t i = (int32_t)(f - f0);
            uint32_t nxt1[4] = {0, 0, 0, 0}, nxt2[4] = {0, 0, 0, 0};
            if (f + 2 < f1) {
                mf_load(a.raw1 + (f + 2) * bpf1, nw1, l, nxt1);
                mf_load(a.raw2 + (f + 2) * bpf2, nw2, l, nxt2);
            }
            const int s1 = __builtin_amdgcn_readlane(pf.s1, jf);
            const int s2 = __builtin_amdgcn_readlane(pf.s2, jf);
            // stage 1 of both stations by column half: each B operand read
            // from LDS feeds two MFMAs, each twiddle load two stations
            half8 a10, a11, a20, a21;
            mf_build_a(row1, s1, l, a10, a11);
            mf_build_a(row2, s2, l, a20, a21);
#pragma unroll 1
            for (int h = 0; h < 2; ++h) {
                f4 y1[2], y2[2];   // tiles h (re) and 2 + h (im)
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const int t = h + 2 * q;
                    const half8 bh0 = ld_b1(sh.b1, t * 2 + 0, l), bl0 = ld_b1(sh.b1, t * 2 + 1, l);
                    const half8 bh1 = ld_b1(sh.b1, (4 + t) * 2 + 0, l), bl1 = ld_b1(sh.b1, (4 + t) * 2 + 1, l);
                    y1[q] = mf_stage1(a10, a11, bh0, bl0, bh1, bl1);
                    y2[q] = mf_stage1(a20, a21, bh0, bl0, bh1, bl1);
                }
                float2 tw[4];
                mf_load_tw(sh.tw2, h, l, tw);
                // the slot was last read by the consumer for frame i - 1
                if (h == 0) mf_wait_ge(&sh.cons[p], i, ok, &sh.broken);
                mf_twiddle_store_half(y1[0], y1[1], tw, &sh.z[p][0][0], 0, h, l);
                mf_twiddle_store_half(y2[0], y2[1], tw, &sh.z[p][0][0], 1, h, l);
            }
            wave_lds_sync();
            mf_columns_dft16(&sh.z[p][0][0], l);
            mf_publish(&sh.prod[p], i + 1);
            if (f + 1 < f1) {   // decode frame f+1 (the staging rows are this wave's own)
                wave_lds_sync();
                mf_stage(cur1, bits1, kind1, nw1, sh.slut[0], sh.dlut[0], row1 + mf::PAD, l);
                mf_stage(cur2, bits2, kind2, nw2, sh.slut[1], sh.dlut[1], row2 + mf::PAD, l);
            }
            wave_lds_sync();
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                cur1[u] = nxt1[u];
                cur2[u] = nxt2[u];
            }
        }
    }
}

// Products of one frame from Z/2 in LDS (as k_sector_mf2's epilogue).
// FULL: the overlap is bins 0..M-1 and every ACF bin is wanted, so the only
// bin outside the overlap is M (lane 0's slot 1): per-lane factors k11 / k22
// / kM (0 there, else 1) replace the slot masks.
// M: points per frame, NP: bin pairs per lane (p = l + 64 j, j < NP).
template <bool FULL, int M = mf::M, int NP = 4>
__device__ __forceinline__ void mf_mac(const float2 *z1, const float2 *z2, const float *twu,
                                       const MfFrame &pf, int jf, int l, bool do_acf, bool do_xcf,
                                       uint32_t m11, uint32_t m22, uint32_t m_ov, float k11,
                                       float k22, float kM, double (&acc)[2 * NP][4]) {
    const double phw = readlane_d(pf.phw, jf);
    const double dth = readlane_d(pf.dth, jf);
    const float2 F1 = make_float2(readlane_f(pf.F1.x, jf), readlane_f(pf.F1.y, jf));
    const float2 Cm = make_float2(readlane_f(pf.C.x, jf), readlane_f(pf.C.y, jf));
    const float r0 = (float)wrap_turns(fma((double)l, dth, phw));
    float2 P = make_float2(__builtin_amdgcn_cosf(r0), __builtin_amdgcn_sinf(r0));
#pragma unroll
    for (int j = 0; j < NP; ++j) {
        const int p = l + 64 * j;
        const int pm = (M - p) & (M - 1);
        const float2 w = make_float2(twu[(2 * j) * 64 + l], twu[(2 * j + 1) * 64 + l]);
        float2 x1p, x1m, x2p, x2m;
        untangle_pair(z1[p], z1[pm], w, x1p, x1m);
        untangle_pair(z2[p], z2[pm], w, x2p, x2m);
        const int sp = 2 * j, sm = 2 * j + 1;
        if (do_acf) {
            if (FULL) {
                acc[sp][0] += (double)norm_sqr(x1p);
                acc[sp][1] += (double)norm_sqr(x2p);
                acc[sm][0] += (double)(j == 0 ? k11 * norm_sqr(x1m) : norm_sqr(x1m));
                acc[sm][1] += (double)(j == 0 ? k22 * norm_sqr(x2m) : norm_sqr(x2m));
            } else {
                acc[sp][0] += (double)(((m11 >> sp) & 1u) ? norm_sqr(x1p) : 0.0f);
                acc[sp][1] += (double)(((m22 >> sp) & 1u) ? norm_sqr(x2p) : 0.0f);
                acc[sm][0] += (double)(((m11 >> sm) & 1u) ? norm_sqr(x1m) : 0.0f);
                acc[sm][1] += (double)(((m22 >> sm) & 1u) ? norm_sqr(x2m) : 0.0f);
            }
        }
        if (do_xcf) {
            float2 Pp, Pm;
            if (FULL) {
                Pp = P;
                Pm = c_mul(Cm, c_conj(P));
                if (j == 0) Pm = make_float2(kM * Pm.x, kM * Pm.y);
            } else {
                Pp = ((m_ov >> sp) & 1u) ? P : make_float2(0.f, 0.f);
                Pm = ((m_ov >> sm) & 1u) ? c_mul(Cm, c_conj(P)) : make_float2(0.f, 0.f);
            }
            const float2 vp = c_mul(c_mul(x1p, c_conj(x2p)), Pp);
            const float2 vm = c_mul(c_mul(x1m, c_conj(x2m)), Pm);
            acc[sp][2] += (double)vp.x;
            acc[sp][3] += (double)vp.y;
            acc[sm][2] += (double)vm.x;
            acc[sm][3] += (double)vm.y;
        }
        P = c_mul(P, F1);
    }
}

// The same products accumulated in FP32 over a batch of frames, each
// product folded into its accumulator by fused multiply-adds (|x|^2 =
// fma(x.x, x.x, fma(x.y, x.y, acc)); (u P).re / .im likewise).  The batch
// sums go to FP64 totals once per batch (mf_flush).
template <bool FULL>
__device__ __forceinline__ void mf_mac32(const float2 *z1, const float2 *z2, const float *twu,
                                         const MfFrame &pf, int jf, int l, bool do_acf, bool do_xcf,
                                         uint32_t m11, uint32_t m22, uint32_t m_ov, float k11,
                                         float k22, float kM, float (&acc)[8][4]) {
    constexpr int M = mf::M;
    const double phw = readlane_d(pf.phw, jf);
    const double dth = readlane_d(pf.dth, jf);
    const float2 F1 = make_float2(readlane_f(pf.F1.x, jf), readlane_f(pf.F1.y, jf));
    const float2 Cm = make_float2(readlane_f(pf.C.x, jf), readlane_f(pf.C.y, jf));
    const float r0 = (float)wrap_turns(fma((double)l, dth, phw));
    float2 P = make_float2(__builtin_amdgcn_cosf(r0), __builtin_amdgcn_sinf(r0));
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int p = l + 64 * j;
        const int pm = (M - p) & (M - 1);
        const float2 w = make_float2(twu[(2 * j) * 64 + l], twu[(2 * j + 1) * 64 + l]);
        float2 x1p, x1m, x2p, x2m;
        untangle_pair(z1[p], z1[pm], w, x1p, x1m);
        untangle_pair(z2[p], z2[pm], w, x2p, x2m);
        const int sp = 2 * j, sm = 2 * j + 1;
        if (do_acf) {
            float2 y1m = x1m, y2m = x2m;
            bool a1p = true, a2p = true, a1m = true, a2m = true;
            if (FULL) {
                if (j == 0) {   // bin M (lane 0) on the grid of the other station
                    y1m = make_float2(k11 * x1m.x, k11 * x1m.y);
                    y2m = make_float2(k22 * x2m.x, k22 * x2m.y);
                }
            } else {
                a1p = (m11 >> sp) & 1u;
                a2p = (m22 >> sp) & 1u;
                a1m = (m11 >> sm) & 1u;
                a2m = (m22 >> sm) & 1u;
            }
            const float n1p = fmaf(x1p.x, x1p.x, fmaf(x1p.y, x1p.y, acc[sp][0]));
            const float n2p = fmaf(x2p.x, x2p.x, fmaf(x2p.y, x2p.y, acc[sp][1]));
            const float n1m = fmaf(y1m.x, x1m.x, fmaf(y1m.y, x1m.y, acc[sm][0]));
            const float n2m = fmaf(y2m.x, x2m.x, fmaf(y2m.y, x2m.y, acc[sm][1]));
            acc[sp][0] = a1p ? n1p : acc[sp][0];
            acc[sp][1] = a2p ? n2p : acc[sp][1];
            acc[sm][0] = a1m ? n1m : acc[sm][0];
            acc[sm][1] = a2m ? n2m : acc[sm][1];
        }
        if (do_xcf) {
            float2 Pp, Pm;
            if (FULL) {
                Pp = P;
                Pm = c_mul(Cm, c_conj(P));
                if (j == 0) Pm = make_float2(kM * Pm.x, kM * Pm.y);
            } else {
                Pp = ((m_ov >> sp) & 1u) ? P : make_float2(0.f, 0.f);
                Pm = ((m_ov >> sm) & 1u) ? c_mul(Cm, c_conj(P)) : make_float2(0.f, 0.f);
            }
            const float2 up = c_mul(x1p, c_conj(x2p));
            const float2 um = c_mul(x1m, c_conj(x2m));
            acc[sp][2] = fmaf(up.x, Pp.x, fmaf(-up.y, Pp.y, acc[sp][2]));
            acc[sp][3] = fmaf(up.x, Pp.y, fmaf(up.y, Pp.x, acc[sp][3]));
            acc[sm][2] = fmaf(um.x, Pm.x, fmaf(-um.y, Pm.y, acc[sm][2]));
            acc[sm][3] = fmaf(um.x, Pm.y, fmaf(um.y, Pm.x, acc[sm][3]));
        }
        P = c_mul(P, F1);
    }
}

// This consumer lane's FP64 totals: 32 values as [16 pairs][64 lanes]
// double2 in global scratch (each wave access is one coalesced 1 KB line
// set; the 16 KB per consumer stays in L2).  Adds the batch's FP32 sums in
// a fixed order (batch after batch) and clears them; `first` writes
// instead of adding.
__device__ __forceinline__ void mf_flush(double *g, float (&acc)[8][4], bool first, int l) {
    double2 *g2 = reinterpret_cast<double2 *>(g) + l;
#pragma unroll
    for (int v0 = 0; v0 < 16; v0 += 4) {   // four 16-byte values in flight at a time
        double2 t[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) t[u] = first ? make_double2(0.0, 0.0) : g2[(v0 + u) * 64];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int v = v0 + u, s = v >> 1, c = 2 * (v & 1);
            t[u].x += (double)acc[s][c];
            t[u].y += (double)acc[s][c + 1];
            g2[v * 64] = t[u];
            acc[s][c] = 0.0f;
            acc[s][c + 1] = 0.0f;
        }
    }
}

__device__ __forceinline__ void mf_load_totals(const double *g, bool none, double (&acc)[8][4], int l) {
    const double2 *g2 = reinterpret_cast<const double2 *>(g);
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const double2 t = none ? make_double2(0.0, 0.0) : g2[v * 64 + l];
        acc[v >> 1][2 * (v & 1)] = t.x;
        acc[v >> 1][2 * (v & 1) + 1] = t.y;
    }
}

// Bin k = M/2 pairs with itself: X = 2 conj(Z'[M/2]).  Its products are
// formed once per batch of frames, lane l taking the frame whose
// parameters it holds (u1 / u2: that frame's Z'[M/2]), and added to this
// lane's FP64 sums; the lanes' sums are added at the end of the kernel.
__device__ __forceinline__ void mf_mid(const MfFrame &pf, float2 u1, float2 u2, bool do_acf,
                                       bool do_xcf, bool in11, bool in22, bool in_ov,
                                       double (&macc)[4], int kmid = mf::M / 2) {
    const float r = (float)wrap_turns(fma((double)kmid, pf.dth, pf.phw));
    const float2 P = make_float2(__builtin_amdgcn_cosf(r), __builtin_amdgcn_sinf(r));
    const float2 x1 = make_float2(2.0f * u1.x, -2.0f * u1.y);
    const float2 x2 = make_float2(2.0f * u2.x, -2.0f * u2.y);
    if (do_acf) {
        macc[0] += (double)(in11 ? norm_sqr(x1) : 0.0f);
        macc[1] += (double)(in22 ? norm_sqr(x2) : 0.0f);
    }
    if (do_xcf) {
        const float2 v = c_mul(c_mul(x1, c_conj(x2)), in_ov ? P : make_float2(0.f, 0.f));
        macc[2] += (double)v.x;
        macc[3] += (double)v.y;
    }
}

// Consumer c: frames [fa0, fa1) of producer c and [fb0, fb1) of producer
// c + 4, alternating (frame i of c, frame i of c + 4, frame i + 1 of c, ...).
// ACC32: each 64-iteration batch (128 frames) is summed in FP32 registers
// (mf_mac32) and added to FP64 totals in global scratch at the batch's end;
// otherwise every frame's products are added to FP64 registers (mf_mac).
template <bool FULL, bool ACC32>
__device__ void mf3_consumer(const SectorArgs &a, MfShared3 &sh, int c, int l, int64_t fa0,
                             int64_t fa1, int64_t fb0, int64_t fb1, double (&acc)[8][4],
                             double (&macc)[4]) {
    const bool do_acf = (a.flags & FLAG_ACF) != 0;
    const bool do_xcf = (a.flags & FLAG_XCF) != 0;
    const bool ov_only = (a.flags & FLAG_ACF_OVERLAP) != 0;
    const int gs = a.a1s, ge = a.a1s + a.ov;   // a1s == a2s on this path
    uint32_t m_grid = 0, m_ov = 0;
#pragma unroll
    for (int s = 0; s < 9; ++s) {
        const int k = mf_bin(s, l);
        const bool in_ov = k >= gs && k < ge;
        if (!ov_only || in_ov) m_grid |= 1u << s;
        if (in_ov) m_ov |= 1u << s;
    }
    const uint32_t m11 = a.grid == 0 ? m_grid : m_ov;
    const uint32_t m22 = a.grid == 0 ? m_ov : m_grid;
    const float kM = (l == 0) ? 0.0f : 1.0f;   // FULL: bin M (lane 0, slot 1) is outside the overlap
    const float k11 = a.grid == 0 ? 1.0f : kM, k22 = a.grid == 0 ? kM : 1.0f;
    const bool in11 = (m11 >> 8) & 1u, in22 = (m22 >> 8) & 1u, in_ov = (m_ov >> 8) & 1u;  // bin M/2
    const int64_t na = fa1 - fa0, nb = fb1 - fb0;
    const int64_t n = na > nb ? na : nb;
    const int q = c + mf::CONS;
    bool ok = true;
    double *gacc = ACC32 ? a.cons_acc + ((size_t)blockIdx.x * mf::CONS + c) * 32 * 64 : nullptr;
    float acc32[8][4];
#pragma unroll
    for (int s = 0; s < 8; ++s)
#pragma unroll
        for (int k = 0; k < 4; ++k) acc32[s][k] = 0.0f;
    bool first = true;
    for (int64_t g = 0; g < n; g += mf::FB) {
        const int cnt = (int)((n - g) < mf::FB ? (n - g) : mf::FB);
        MfFrame pa = {}, pb = {};
        if (l < cnt && g + l < na) pa = mf_frame(a.first_frame + fa0 + g + l, a);
        if (l < cnt && g + l < nb) pb = mf_frame(a.first_frame + fb0 + g + l, a);
        float2 ua1 = make_float2(0.f, 0.f), ua2 = ua1, ub1 = ua1, ub2 = ua1;   // Z'[M/2] of frame g + l
        for (int jf = 0; jf < cnt; ++jf) {
            const int32_t i = (int32_t)(g + jf);
            if (i < na) {
                mf_wait_ge(&sh.prod[c], i + 1, ok, &sh.broken);
                if constexpr (ACC32)
                    mf_mac32<FULL>(sh.z[c][0], sh.z[c][1], sh.twu, pa, jf, l, do_acf, do_xcf, m11,
                                   m22, m_ov, k11, k22, kM, acc32);
                else
                    mf_mac<FULL>(sh.z[c][0], sh.z[c][1], sh.twu, pa, jf, l, do_acf, do_xcf, m11, m22,
                                 m_ov, k11, k22, kM, acc);
                const float2 u1 = sh.z[c][0][mf::M / 2], u2 = sh.z[c][1][mf::M / 2];
                if (l == jf) {
                    ua1 = u1;
                    ua2 = u2;
                }
                mf_publish(&sh.cons[c], i + 1);   // this slot's reads are done
            }
            if (i < nb) {
                mf_wait_ge(&sh.prod[q], i + 1, ok, &sh.broken);
                if constexpr (ACC32)
                    mf_mac32<FULL>(sh.z[q][0], sh.z[q][1], sh.twu, pb, jf, l, do_acf, do_xcf, m11,
                                   m22, m_ov, k11, k22, kM, acc32);
                else
                    mf_mac<FULL>(sh.z[q][0], sh.z[q][1], sh.twu, pb, jf, l, do_acf, do_xcf, m11, m22,
                                 m_ov, k11, k22, kM, acc);
                const float2 u1 = sh.z[q][0][mf::M / 2], u2 = sh.z[q][1][mf::M / 2];
                if (l == jf) {
                    ub1 = u1;
                    ub2 = u2;
                }
                mf_publish(&sh.cons[q], i + 1);
            }
        }
        if (l < cnt && g + l < na) mf_mid(pa, ua1, ua2, do_acf, do_xcf, in11, in22, in_ov, macc);
        if (l < cnt && g + l < nb) mf_mid(pb, ub1, ub2, do_acf, do_xcf, in11, in22, in_ov, macc);
        if constexpr (ACC32) mf_flush(gacc, acc32, first, l);
        first = false;
    }
    if constexpr (ACC32) mf_load_totals(gacc, first, acc, l);
}

// The producer/consumer kernels' LDS set-up: operand, twiddle and decode
// tables from the plan's MFT_* table, zeroed staging rows, hand-over flags.
__device__ __forceinline__ void mf3_setup(const SectorArgs &a, MfShared3 &sh, int tid) {
    constexpr int NT = mf::THREADS3;
    const uint8_t *tab = reinterpret_cast<const uint8_t *>(a.mf_tab);
    {
        const uint4 *g = reinterpret_cast<const uint4 *>(tab + MFT_B1);
        for (int i = tid; i < 16 * 64; i += NT) sh.b1[i] = g[i];
        const uint32_t *gs = reinterpret_cast<const uint32_t *>(tab + MFT_SLUT);
        for (int i = tid; i < 2 * 4 * 256; i += NT) (&sh.slut[0][0])[i] = gs[i];
        const uint32_t *gd = reinterpret_cast<const uint32_t *>(tab + MFT_DLUT);
        for (int i = tid; i < 2 * MFT_DLUT_HALVES / 2; i += NT)
            reinterpret_cast<uint32_t *>(&sh.dlut[0][0])[i] = gd[i];
        uint32_t *st32 = reinterpret_cast<uint32_t *>(&sh.stage[0][0][0]);
        for (int i = tid; i < mf::PROD * 2 * mf::SROW3 / 2; i += NT) st32[i] = 0u;
        const float *ga = reinterpret_cast<const float *>(tab + MFT_A2);
        const float *gt = reinterpret_cast<const float *>(tab + MFT_TW2);
        const float *gu = reinterpret_cast<const float *>(tab + MFT_TWU);
        for (int i = tid; i < 16 * 64; i += NT) {
            sh.a2[i] = ga[i];
            sh.tw2[i] = gt[i];
        }
        for (int i = tid; i < 8 * 64; i += NT) sh.twu[i] = gu[i];
        if (tid < mf::PROD) {
            sh.prod[tid] = 0;
            sh.cons[tid] = 0;
        }
        if (tid == 0) sh.broken = 0;
    }
}

template <bool ACC32>
__global__ __launch_bounds__(mf::THREADS3) void k_sector_mf3(SectorArgs a) {
    __shared__ MfShared3 sh;
    constexpr int M = mf::M;
    constexpr int NT = mf::THREADS3;
    const int tid = threadIdx.x;
    const int l = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    mf3_setup(a, sh, tid);
    __syncthreads();

    // frames of producer p: a contiguous eighth of the workgroup's range
    const int64_t wg0 = (int64_t)blockIdx.x * a.frames_per_wg;
    int64_t wg1 = wg0 + a.frames_per_wg;
    if (wg1 > a.nf) wg1 = a.nf;
    const int64_t per = (a.frames_per_wg + mf::PROD - 1) / mf::PROD;
    auto range = [&](int p, int64_t &r0, int64_t &r1) {
        r0 = wg0 + per * p;
        if (r0 > wg1) r0 = wg1;
        r1 = (r0 + per < wg1) ? r0 + per : wg1;
    };

    // consumers: FP64 totals of their bins (ACC32: loaded from the batch
    // totals when the consumer finishes)
    double acc[8][4], macc[4] = {0.0, 0.0, 0.0, 0.0};
    if constexpr (!ACC32) {
#pragma unroll
        for (int s = 0; s < 8; ++s)
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[s][c] = 0.0;
    }
    if (wv < mf::PROD) {
        int64_t f0, f1;
        range(wv, f0, f1);
        mf3_producer(a, sh, wv, l, f0, f1);
    } else {
        const int c = wv - mf::PROD;
        int64_t fa0, fa1, fb0, fb1;
        range(c, fa0, fa1);
        range(c + mf::CONS, fb0, fb1);
        const bool full = (a.flags & FLAG_ACF_OVERLAP) == 0 && a.a1s == 0 && a.ov == M;
        if (full) mf3_consumer<true, ACC32>(a, sh, c, l, fa0, fa1, fb0, fb1, acc, macc);
        else mf3_consumer<false, ACC32>(a, sh, c, l, fa0, fa1, fb0, fb1, acc, macc);
    }

    // ---- workgroup partial: consumers add their bins in a fixed order
    __syncthreads();
    double *red = reinterpret_cast<double *>(&sh.z[0][0][0]);   // [4][M + 1]
    double *mids = reinterpret_cast<double *>(&sh.z[4][0][0]);  // [CONS][4][64] lanes' bin-M/2 sums
    constexpr int H = M + 1;
    if (wv >= mf::PROD)
#pragma unroll
        for (int c = 0; c < 4; ++c) mids[((wv - mf::PROD) * 4 + c) * 64 + l] = macc[c];
    __syncthreads();
    for (int w = mf::PROD; w < mf::PROD + mf::CONS; ++w) {
        if (wv == w) {
            const bool first = w == mf::PROD;
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const int k = mf_bin(s, l);
#pragma unroll
                for (int c = 0; c < 4; ++c) red[c * H + k] = (first ? 0.0 : red[c * H + k]) + acc[s][c];
            }
            if (l == 0)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const double *mc = mids + ((w - mf::PROD) * 4 + c) * 64;
                    double sum = 0.0;
                    for (int i = 0; i < 64; ++i) sum += mc[i];   // lanes in order
                    red[c * H + M / 2] = (first ? 0.0 : red[c * H + M / 2]) + sum;
                }
        }
        __syncthreads();
    }
    double *out = a.part + (size_t)blockIdx.x * kNumProducts * H;
    const double bad = sh.broken ? __builtin_nan("") : 0.0;   // NaN marks a failed hand-over
    if (sh.broken && tid == 0 && a.err)   // and the plan's error word reports it to the host
        __hip_atomic_fetch_or(a.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int k = tid; k < H; k += NT) {
        out[P_PH * H + k] = bad;
        out[P_11 * H + k] = red[0 * H + k] + bad;
        out[P_22 * H + k] = red[1 * H + k] + bad;
        out[P_12R * H + k] = red[2 * H + k] + bad;
        out[P_12I * H + k] = red[3 * H + k] + bad;
    }
}

// ------------------------------- N = 256 / 512: packed station-frames
// The producer/consumer kernel for FFT lengths N' = 1024 / FPS (FPS = 4:
// N' = 256; FPS = 2: N' = 512).  A producer slot holds FPS consecutive
// frames of each station, whose M' = 32 R2 points (R2 = 16 / FPS, m = R2 m1
// + m2) fill the 16 rows of one DFT-32 tile: row rho = R2 q + m2 for frame
// q.  Stage 1 is k_sector_mf3's (the same operands and MFMAs); the plan's
// twiddle table holds W_M'^{m2 k1}; stage 2 is a DFT-R2 over m2 per frame in
// each column lane, which writes Z'/2 of frame q at slot offset q M'.  The
// consumers take the frames of a slot in turn: lane l the bin pairs
// (p, M' - p), p = l + 64 j (j < M' / 128), with k_sector_mf3's products,
// and bin M'/2 once per 64-frame batch.
namespace mfs {
template <int FPS>
struct Shape {
    static constexpr int NF = mf::N / FPS;     // samples per frame
    static constexpr int MF = mf::M / FPS;     // complex points per frame
    static constexpr int R2 = 16 / FPS;        // rows m2 per frame
    static constexpr int PAD = 32;             // zero samples each side of a staged frame
    static constexpr int SUB = NF + 2 * PAD;   // halves per staged frame
    static constexpr int NP = MF / 128;        // bin pairs per lane per frame
    static_assert(FPS * SUB <= mf::SROW3, "staging row");
};
}  // namespace mfs

// Bins of lane l: slot 2j -> p = l + 64 j, slot 2j + 1 -> M' - p; slot 2 NP
// -> M'/2.
template <int FPS>
__device__ __forceinline__ int mfs_bin(int slot, int l) {
    using S = mfs::Shape<FPS>;
    if (slot == 2 * S::NP) return S::MF / 2;
    const int p = l + 64 * (slot >> 1);
    return (slot & 1) ? S::MF - p : p;
}

// A operands of a packed slot: lane l supplies row rho = l % 16 (frame q =
// rho / R2, m2 = rho % R2) with that frame's shift sh: sample 2(R2 m1 + m2)
// + e of frame q, m1 = 16 kap + g + 4 j, staged at q SUB + PAD + 2(R2 m1 +
// m2) - sh.  Shifts within the padding read each FP16 pair by a funnel shift
// of two dwords; a lane with a larger shift checks every sample against its
// frame.
template <int FPS>
__device__ __forceinline__ void mfs_build_a(const uint16_t *row, int sh, int l, half8 &a0, half8 &a1) {
    using S = mfs::Shape<FPS>;
    constexpr int R2 = S::R2, OK = 16 * R2, OJ = 4 * R2;   // dword offsets of K-step kap and slot pair j
    const int rho = l & 15, q = rho / R2, m2 = rho % R2;
    uint32_t d[8];
    if (sh >= -S::PAD && sh <= S::PAD) {
        const int base = q * S::SUB + S::PAD + 2 * m2 - sh + 2 * R2 * (l >> 4);   // halves, >= 0
        const uint32_t *p0 = reinterpret_cast<const uint32_t *>(row) + (base >> 1);
        const uint32_t bs = (uint32_t)(base & 1) * 2u;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int o = OK * (k >> 2) + OJ * (k & 3);
            d[k] = __builtin_amdgcn_alignbyte(p0[o + 1], p0[o], bs);
        }
    } else {
        const uint16_t *fr = row + q * S::SUB + S::PAD;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int m1 = 16 * (k >> 2) + (l >> 4) + 4 * (k & 3);
            const int s0 = 2 * (R2 * m1 + m2) - sh;
            const uint32_t lo = (unsigned)s0 < (unsigned)S::NF ? (uint32_t)fr[s0] : 0u;
            const uint32_t hi = (unsigned)(s0 + 1) < (unsigned)S::NF ? (uint32_t)fr[s0 + 1] : 0u;
            d[k] = lo | (hi << 16);
        }
    }
    a0 = __builtin_bit_cast(half8, make_uint4(d[0], d[1], d[2], d[3]));
    a1 = __builtin_bit_cast(half8, make_uint4(d[4], d[5], d[6], d[7]));
}

// DFT-8 of x[0..7] (radix 2 x 4: even and odd DFT-4s, W8 twiddles).
__device__ __forceinline__ void mf_dft8(const float2 (&x)[8], float2 (&o)[8]) {
    float2 e[4], d[4];
    mf_dft4(x[0], x[2], x[4], x[6], e[0], e[1], e[2], e[3]);
    mf_dft4(x[1], x[3], x[5], x[7], d[0], d[1], d[2], d[3]);
    constexpr float c = 0.70710678118654752f;
    const float2 t1 = make_float2(c * (d[1].x + d[1].y), c * (d[1].y - d[1].x));    // W8^1 = (1 - i)/sqrt2
    const float2 t2 = make_float2(d[2].y, -d[2].x);                                 // W8^2 = -i
    const float2 t3 = make_float2(c * (d[3].y - d[3].x), -c * (d[3].x + d[3].y));   // W8^3 = -(1 + i)/sqrt2
    const float2 t[4] = {d[0], t1, t2, t3};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        o[k] = make_float2(e[k].x + t[k].x, e[k].y + t[k].y);
        o[k + 4] = make_float2(e[k].x - t[k].x, e[k].y - t[k].y);
    }
}

// Stage 2 of a packed slot: lane l takes column k1 = l & 31 of station
// l >> 5 (its 16 rows from the transposed copy) and runs a DFT-R2 over m2
// for each frame q, writing Z'/2[q M' + k1 + 32 k2].
template <int FPS>
__device__ __forceinline__ void mfs_columns(float2 *slot, int l) {
    using S = mfs::Shape<FPS>;
    const int s = l >> 5, k1 = l & 31;
    float2 x[16];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const float4 v = *reinterpret_cast<const float4 *>(slot + mf_tidx(s, k1, j));
        x[2 * j] = make_float2(v.x, v.y);
        x[2 * j + 1] = make_float2(v.z, v.w);
    }
    wave_lds_sync();   // every lane's column is read before Z overwrites the slot
    float2 *zo = slot + s * mf::M + k1;
#pragma unroll
    for (int q = 0; q < FPS; ++q) {
        if constexpr (S::R2 == 4) {
            float2 o0, o1, o2, o3;
            mf_dft4(x[4 * q], x[4 * q + 1], x[4 * q + 2], x[4 * q + 3], o0, o1, o2, o3);
            zo[q * S::MF] = o0;
            zo[q * S::MF + 32] = o1;
            zo[q * S::MF + 64] = o2;
            zo[q * S::MF + 96] = o3;
        } else {
            const float2 xi[8] = {x[8 * q], x[8 * q + 1], x[8 * q + 2], x[8 * q + 3],
                                  x[8 * q + 4], x[8 * q + 5], x[8 * q + 6], x[8 * q + 7]};
            float2 o[8];
            mf_dft8(xi, o);
#pragma unroll
            for (int k2 = 0; k2 < 8; ++k2) zo[q * S::MF + 32 * k2] = o[k2];
        }
    }
}

// Producer p of the packed kernel: frames [f0, f1) in slots of FPS frames
// (the last slot may hold fewer); the next slot's words are decoded into
// the staging rows while this one is transformed, as in mf3_producer.
template <int FPS>
__device__ void mfs_producer(const SectorArgs &a, MfShared3 &sh, int p, int l, int64_t f0, int64_t f1) {
    using S = mfs::Shape<FPS>;
    const int bits1 = a.st[0].bits, bits2 = a.st[1].bits;
    const int kind1 = a.st[0].shuffle_kind, kind2 = a.st[1].shuffle_kind;
    const int nw1 = a.st[0].nwords, nw2 = a.st[1].nwords;   // words per frame (powers of two)
    const int lg1 = __builtin_ctz((unsigned)nw1), lg2 = __builtin_ctz((unsigned)nw2);
    const int64_t bpf1 = a.st[0].bpf, bpf2 = a.st[1].bpf;
    uint16_t *row1 = sh.stage[p][0];
    uint16_t *row2 = sh.stage[p][1];
    const int64_t n = f1 - f0;
    const int nslots = (int)((n + FPS - 1) / FPS);
    auto nvalid = [&](int i) { return (int)((n - (int64_t)FPS * i) < FPS ? (n - (int64_t)FPS * i) : FPS); };
    auto load = [&](int i, uint32_t (&w1)[4], uint32_t (&w2)[4]) {
        const int64_t fs = f0 + (int64_t)FPS * i;
        const int nv = nvalid(i);
        mf_load(a.raw1 + fs * bpf1, nv * nw1, l, w1);
        mf_load(a.raw2 + fs * bpf2, nv * nw2, l, w2);
    };
    auto stage = [&](int i, const uint32_t (&w1)[4], const uint32_t (&w2)[4]) {
        const int nv = nvalid(i);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int wi = l + 64 * u;
            if (wi < nv * nw1) {
                const int q = wi >> lg1;
                mf_stage_word(w1[u], bits1, kind1, sh.slut[0], sh.dlut[0], row1 + q * S::SUB + S::PAD,
                              wi - (q << lg1));
            }
            if (wi < nv * nw2) {
                const int q = wi >> lg2;
                mf_stage_word(w2[u], bits2, kind2, sh.slut[1], sh.dlut[1], row2 + q * S::SUB + S::PAD,
                              wi - (q << lg2));
            }
        }
    };
    uint32_t cur1[4] = {0, 0, 0, 0}, cur2[4] = {0, 0, 0, 0};
    bool ok = true;
    if (nslots > 0) {
        load(0, cur1, cur2);
        stage(0, cur1, cur2);
        if (nslots > 1) load(1, cur1, cur2);
        wave_lds_sync();
    }
    const int qf = (l & 15) / S::R2;   // this lane's frame within a slot
    for (int64_t fb = f0; fb < f1; fb += mf::FB) {
        const int nb = (int)((f1 - fb) < mf::FB ? (f1 - fb) : mf::FB);
        MfFrame pf = {};
        if (l < nb) pf = mf_frame(a.first_frame + fb + l, a);
        const int ns = (nb + FPS - 1) / FPS;
        for (int js = 0; js < ns; ++js) {
            const int i = (int)((fb - f0) / FPS) + js;
            uint32_t nxt1[4] = {0, 0, 0, 0}, nxt2[4] = {0, 0, 0, 0};
            if (i + 2 < nslots) load(i + 2, nxt1, nxt2);
            int s1 = 0, s2 = 0;
#pragma unroll
            for (int q = 0; q < FPS; ++q) {
                const int jf = (js * FPS + q) < nb ? js * FPS + q : nb - 1;
                const int t1 = __builtin_amdgcn_readlane(pf.s1, jf);
                const int t2 = __builtin_amdgcn_readlane(pf.s2, jf);
                if (qf == q) {
                    s1 = t1;
                    s2 = t2;
                }
            }
            half8 a10, a11, a20, a21;
            mfs_build_a<FPS>(row1, s1, l, a10, a11);
            mfs_build_a<FPS>(row2, s2, l, a20, a21);
#pragma unroll 1
            for (int h = 0; h < 2; ++h) {
                f4 y1[2], y2[2];   // tiles h (re) and 2 + h (im)
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const int t = h + 2 * q;
                    const half8 bh0 = ld_b1(sh.b1, t * 2 + 0, l), bl0 = ld_b1(sh.b1, t * 2 + 1, l);
                    const half8 bh1 = ld_b1(sh.b1, (4 + t) * 2 + 0, l), bl1 = ld_b1(sh.b1, (4 + t) * 2 + 1, l);
                    y1[q] = mf_stage1(a10, a11, bh0, bl0, bh1, bl1);
                    y2[q] = mf_stage1(a20, a21, bh0, bl0, bh1, bl1);
                }
                float2 tw[4];
                mf_load_tw(sh.tw2, h, l, tw);
                // the slot was last read by the consumer for slot i - 1
                if (h == 0) mf_wait_ge(&sh.cons[p], i, ok, &sh.broken);
                mf_twiddle_store_half(y1[0], y1[1], tw, &sh.z[p][0][0], 0, h, l);
                mf_twiddle_store_half(y2[0], y2[1], tw, &sh.z[p][0][0], 1, h, l);
            }
            wave_lds_sync();
            mfs_columns<FPS>(&sh.z[p][0][0], l);
            mf_publish(&sh.prod[p], i + 1);
            if (i + 1 < nslots) {   // decode slot i + 1 (the staging rows are this wave's own)
                wave_lds_sync();
                stage(i + 1, cur1, cur2);
            }
            wave_lds_sync();
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                cur1[u] = nxt1[u];
                cur2[u] = nxt2[u];
            }
        }
    }
}

// Consumer c: the slots of producers c and c + 4, alternating, each slot's
// frames in turn.
template <int FPS, bool FULL>
__device__ void mfs_consumer(const SectorArgs &a, MfShared3 &sh, int c, int l, int64_t fa0,
                             int64_t fa1, int64_t fb0, int64_t fb1,
                             double (&acc)[2 * mfs::Shape<FPS>::NP][4], double (&macc)[4]) {
    using S = mfs::Shape<FPS>;
    constexpr int NP = S::NP, MF = S::MF;
    const bool do_acf = (a.flags & FLAG_ACF) != 0;
    const bool do_xcf = (a.flags & FLAG_XCF) != 0;
    const bool ov_only = (a.flags & FLAG_ACF_OVERLAP) != 0;
    const int gs = a.a1s, ge = a.a1s + a.ov;   // a1s == a2s on this path
    uint32_t m_grid = 0, m_ov = 0;
#pragma unroll
    for (int s = 0; s <= 2 * NP; ++s) {
        const int k = mfs_bin<FPS>(s, l);
        const bool in_ov = k >= gs && k < ge;
        if (!ov_only || in_ov) m_grid |= 1u << s;
        if (in_ov) m_ov |= 1u << s;
    }
    const uint32_t m11 = a.grid == 0 ? m_grid : m_ov;
    const uint32_t m22 = a.grid == 0 ? m_ov : m_grid;
    const float kM = (l == 0) ? 0.0f : 1.0f;   // FULL: bin M' (lane 0, slot 1) is outside the overlap
    const float k11 = a.grid == 0 ? 1.0f : kM, k22 = a.grid == 0 ? kM : 1.0f;
    const bool in11 = (m11 >> (2 * NP)) & 1u, in22 = (m22 >> (2 * NP)) & 1u;
    const bool mid_ov = (m_ov >> (2 * NP)) & 1u;
    const int64_t na = fa1 - fa0, nb = fb1 - fb0;
    const int64_t n = na > nb ? na : nb;
    const int q2 = c + mf::CONS;
    bool ok = true;
    for (int64_t g = 0; g < n; g += mf::FB) {
        const int cnt = (int)((n - g) < mf::FB ? (n - g) : mf::FB);
        MfFrame pa = {}, pb = {};
        if (l < cnt && g + l < na) pa = mf_frame(a.first_frame + fa0 + g + l, a);
        if (l < cnt && g + l < nb) pb = mf_frame(a.first_frame + fb0 + g + l, a);
        float2 ua1 = make_float2(0.f, 0.f), ua2 = ua1, ub1 = ua1, ub2 = ua1;   // Z'[M'/2] of frame g + l
        const int ns = (cnt + FPS - 1) / FPS;
        for (int js = 0; js < ns; ++js) {
            const int i = (int)(g / FPS) + js;
            const int64_t fi = (int64_t)FPS * i;   // the slot's first frame (offset in each range)
            if (fi < na) {
                mf_wait_ge(&sh.prod[c], i + 1, ok, &sh.broken);
#pragma unroll
                for (int q = 0; q < FPS; ++q) {
                    if (fi + q < na) {
                        const int jf = js * FPS + q;
                        const float2 *z1 = &sh.z[c][0][q * MF], *z2 = &sh.z[c][1][q * MF];
                        mf_mac<FULL, MF, NP>(z1, z2, sh.twu, pa, jf, l, do_acf, do_xcf, m11, m22, m_ov,
                                             k11, k22, kM, acc);
                        const float2 u1 = z1[MF / 2], u2 = z2[MF / 2];
                        if (l == jf) {
                            ua1 = u1;
                            ua2 = u2;
                        }
                    }
                }
                mf_publish(&sh.cons[c], i + 1);   // this slot's reads are done
            }
            if (fi < nb) {
                mf_wait_ge(&sh.prod[q2], i + 1, ok, &sh.broken);
#pragma unroll
                for (int q = 0; q < FPS; ++q) {
                    if (fi + q < nb) {
                        const int jf = js * FPS + q;
                        const float2 *z1 = &sh.z[q2][0][q * MF], *z2 = &sh.z[q2][1][q * MF];
                        mf_mac<FULL, MF, NP>(z1, z2, sh.twu, pb, jf, l, do_acf, do_xcf, m11, m22, m_ov,
                                             k11, k22, kM, acc);
                        const float2 u1 = z1[MF / 2], u2 = z2[MF / 2];
                        if (l == jf) {
                            ub1 = u1;
                            ub2 = u2;
                        }
                    }
                }
                mf_publish(&sh.cons[q2], i + 1);
            }
        }
        if (l < cnt && g + l < na) mf_mid(pa, ua1, ua2, do_acf, do_xcf, in11, in22, mid_ov, macc, MF / 2);
        if (l < cnt && g + l < nb) mf_mid(pb, ub1, ub2, do_acf, do_xcf, in11, in22, mid_ov, macc, MF / 2);
    }
}

template <int FPS>
__global__ __launch_bounds__(mf::THREADS3) void k_sector_mfs(SectorArgs a) {
    __shared__ MfShared3 sh;
    using S = mfs::Shape<FPS>;
    constexpr int M = S::MF, NS = 2 * S::NP;
    constexpr int NT = mf::THREADS3;
    const int tid = threadIdx.x;
    const int l = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    mf3_setup(a, sh, tid);
    __syncthreads();
    const int64_t wg0 = (int64_t)blockIdx.x * a.frames_per_wg;
    int64_t wg1 = wg0 + a.frames_per_wg;
    if (wg1 > a.nf) wg1 = a.nf;
    const int64_t per = (a.frames_per_wg + mf::PROD - 1) / mf::PROD;
    auto range = [&](int p, int64_t &r0, int64_t &r1) {
        r0 = wg0 + per * p;
        if (r0 > wg1) r0 = wg1;
        r1 = (r0 + per < wg1) ? r0 + per : wg1;
    };
    double acc[NS][4], macc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[s][c] = 0.0;
    if (wv < mf::PROD) {
        int64_t f0, f1;
        range(wv, f0, f1);
        mfs_producer<FPS>(a, sh, wv, l, f0, f1);
    } else {
        const int c = wv - mf::PROD;
        int64_t fa0, fa1, fb0, fb1;
        range(c, fa0, fa1);
        range(c + mf::CONS, fb0, fb1);
        const bool full = (a.flags & FLAG_ACF_OVERLAP) == 0 && a.a1s == 0 && a.ov == M;
        if (full) mfs_consumer<FPS, true>(a, sh, c, l, fa0, fa1, fb0, fb1, acc, macc);
        else mfs_consumer<FPS, false>(a, sh, c, l, fa0, fa1, fb0, fb1, acc, macc);
    }

    // ---- workgroup partial: consumers add their bins in a fixed order
    __syncthreads();
    double *red = reinterpret_cast<double *>(&sh.z[0][0][0]);   // [4][M' + 1]
    double *mids = reinterpret_cast<double *>(&sh.z[4][0][0]);  // [CONS][4][64] lanes' bin-M'/2 sums
    constexpr int H = M + 1;
    if (wv >= mf::PROD)
#pragma unroll
        for (int c = 0; c < 4; ++c) mids[((wv - mf::PROD) * 4 + c) * 64 + l] = macc[c];
    __syncthreads();
    for (int w = mf::PROD; w < mf::PROD + mf::CONS; ++w) {
        if (wv == w) {
            const bool first = w == mf::PROD;
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const int k = mfs_bin<FPS>(s, l);
#pragma unroll
                for (int c = 0; c < 4; ++c) red[c * H + k] = (first ? 0.0 : red[c * H + k]) + acc[s][c];
            }
            if (l == 0)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const double *mc = mids + ((w - mf::PROD) * 4 + c) * 64;
                    double sum = 0.0;
                    for (int i = 0; i < 64; ++i) sum += mc[i];   // lanes in order
                    red[c * H + M / 2] = (first ? 0.0 : red[c * H + M / 2]) + sum;
                }
        }
        __syncthreads();
    }
    double *out = a.part + (size_t)blockIdx.x * kNumProducts * H;
    const double bad = sh.broken ? __builtin_nan("") : 0.0;   // NaN marks a failed hand-over
    if (sh.broken && tid == 0 && a.err)
        __hip_atomic_fetch_or(a.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int k = tid; k < H; k += NT) {
        out[P_PH * H + k] = bad;
        out[P_11 * H + k] = red[0 * H + k] + bad;
        out[P_22 * H + k] = red[1 * H + k] + bad;
        out[P_12R * H + k] = red[2 * H + k] + bad;
        out[P_12I * H + k] = red[3 * H + k] + bad;
    }
}

// ------------------------------------ phased sum at N = 256 / 512 / 1024
// The producer/consumer kernel with a phased-array consumer (chosen form;
// main.rs:2264-2302, utils.rs:69-124).  The producers are k_sector_mf3's
// (N = 1024, FPS = 1 frame per slot) or the packed kernel's (N = 256 / 512).
// Per frame and output bin the consumer takes each station's bin through its
// fractional-delay rotation exp(2 pi i b th) (th = -df frac turns per bin,
// from the FP64 delay) and fringe phasor, forms cb = w1 c1 + w2 c2, adds
// |cb|^2 to FP64 sums and, for the .raw output, writes cb (DC and Nyquist
// imaginary parts zeroed) to cb_out[frame][M + 1].  Both stations share the
// band start (a1s == a2s), so a lane's bins are its pairs (p, M - p) on both
// stations, and R(M - p) = R(M) conj R(p).
struct MfPhFrame {
    double thg, tho;         // rotation per bin, turns: grid station, other station
    float2 frg, fro;         // fringe phasors
    float2 Sg, So, Cg, Co;   // exp(2 pi i 64 th), exp(2 pi i M th), M = N / 2 of the plan
};

__device__ __forceinline__ float2 mf_phasor(double turns) {
    const float r = (float)wrap_turns(turns);
    return make_float2(__builtin_amdgcn_cosf(r), __builtin_amdgcn_sinf(r));
}

// The chosen phased parameters of frame idx (make_phased_param's values).
__device__ MfPhFrame mfp_frame(int64_t idx, const SectorArgs &a) {
    double tau1, tau2, tf1, tf2;
    frame_taus(idx, a.dly, tau1, tau2, tf1, tf2);
    int64_t n1, n2;
    double fr1, fr2;
    split_delay(tau1, a.dly.fs, n1, fr1);
    split_delay(tau2, a.dly.fs, n2, fr2);
    const float w1 = a.dly.lo1 == 0.0 ? 0.0f : (float)wrap_turns(-a.dly.lo1 * tf1);
    const float w2 = a.dly.lo2 == 0.0 ? 0.0f : (float)wrap_turns(-a.dly.lo2 * tf2);
    const float2 f1 = make_float2(__builtin_amdgcn_cosf(w1), __builtin_amdgcn_sinf(w1));
    const float2 f2 = make_float2(__builtin_amdgcn_cosf(w2), __builtin_amdgcn_sinf(w2));
    const double th1 = -a.df * fr1, th2 = -a.df * fr2;
    const bool g0 = a.grid == 0;
    MfPhFrame o;
    o.thg = g0 ? th1 : th2;
    o.tho = g0 ? th2 : th1;
    o.frg = g0 ? f1 : f2;
    o.fro = g0 ? f2 : f1;
    o.Sg = mf_phasor(64.0 * o.thg);
    o.So = mf_phasor(64.0 * o.tho);
    o.Cg = mf_phasor((double)a.M * o.thg);
    o.Co = mf_phasor((double)a.M * o.tho);
    return o;
}

__device__ __forceinline__ float2 readlane_f2(float2 v, int lane) {
    return make_float2(readlane_f(v.x, lane), readlane_f(v.y, lane));
}

// One output bin (k_phased's arithmetic): the grid station's spectrum value
// xg and the other station's xo on the same bin, their rotations rg / ro;
// nyq zeroes the rotated imaginary parts (bin M, utils.rs:118-123); the
// other station counts only inside the overlap.
__device__ __forceinline__ float2 mfp_combine(float2 xg, float2 rg, float2 xo, float2 ro, bool nyq,
                                              bool in_ov, float2 frg, float2 fro, int gsn, float w1,
                                              float w2) {
    float2 sg = c_mul(xg, rg);
    float2 so = c_mul(xo, ro);
    if (nyq) {
        sg.y = 0.0f;
        so.y = 0.0f;
    }
    const float2 cg = c_mul(sg, frg);
    const float2 co0 = c_mul(so, fro);
    const float2 co = in_ov ? co0 : make_float2(0.0f, 0.0f);
    const float2 c1 = gsn == 0 ? cg : co, c2 = gsn == 0 ? co : cg;
    float2 cb;
    {
#pragma clang fp contract(off)
        const float2 a1 = make_float2(c1.x * w1, c1.y * w1), a2 = make_float2(c2.x * w2, c2.y * w2);
        cb = make_float2(a1.x + a2.x, a1.y + a2.y);
    }
    return cb;
}

// The phased products of one frame of M points (z1 / z2) for lane l's bin
// pairs p = l + 64 j, j < NP.
template <int M, int NP>
__device__ __forceinline__ void mfp_mac(const float2 *z1, const float2 *z2, const float *twu,
                                        const MfPhFrame &pf, int jf, int l, int gsn, uint32_t m_ov,
                                        float w1, float w2, float2 *cbo, double (&acc)[2 * NP]) {
    const double thg = readlane_d(pf.thg, jf), tho = readlane_d(pf.tho, jf);
    const float2 frg = readlane_f2(pf.frg, jf), fro = readlane_f2(pf.fro, jf);
    const float2 Sg = readlane_f2(pf.Sg, jf), So = readlane_f2(pf.So, jf);
    const float2 Cg = readlane_f2(pf.Cg, jf), Co = readlane_f2(pf.Co, jf);
    float2 Rg = mf_phasor((double)l * thg), Ro = mf_phasor((double)l * tho);
    const float2 *zg = gsn == 0 ? z1 : z2;
    const float2 *zo = gsn == 0 ? z2 : z1;
#pragma unroll
    for (int j = 0; j < NP; ++j) {
        const int p = l + 64 * j;
        const int pm = (M - p) & (M - 1);
        const float2 w = make_float2(twu[(2 * j) * 64 + l], twu[(2 * j + 1) * 64 + l]);
        float2 xgp, xgm, xop, xom;
        untangle_pair(zg[p], zg[pm], w, xgp, xgm);
        untangle_pair(zo[p], zo[pm], w, xop, xom);
        const int sp = 2 * j, sm = 2 * j + 1;
        const bool nyq = j == 0 && l == 0;   // lane 0's slot 1 is bin M
        const float2 cbp = mfp_combine(xgp, Rg, xop, Ro, false, (m_ov >> sp) & 1u, frg, fro, gsn, w1, w2);
        const float2 cbm = mfp_combine(xgm, c_mul(Cg, c_conj(Rg)), xom, c_mul(Co, c_conj(Ro)), nyq,
                                       (m_ov >> sm) & 1u, frg, fro, gsn, w1, w2);
        acc[sp] += (double)norm_sqr(cbp);
        acc[sm] += (double)norm_sqr(cbm);
        if (cbo != nullptr) {   // main.rs:2283-2284: DC and Nyquist imaginary parts zeroed
            cbo[p] = make_float2(cbp.x, p == 0 ? 0.0f : cbp.y);
            cbo[M - p] = make_float2(cbm.x, nyq ? 0.0f : cbm.y);
        }
        Rg = c_mul(Rg, Sg);
        Ro = c_mul(Ro, So);
    }
}

// Bin K = M/2 (X = 2 conj Z'[M/2] on each station) of the frame whose
// parameters this lane holds, formed once per batch as in mf_mid.
__device__ __forceinline__ void mfp_mid(const MfPhFrame &pf, float2 u1, float2 u2, int gsn, bool in_ov,
                                        float w1, float w2, float2 *cb_mid, double &macc, int K) {
    const float2 x1 = make_float2(2.0f * u1.x, -2.0f * u1.y);
    const float2 x2 = make_float2(2.0f * u2.x, -2.0f * u2.y);
    const float2 xg = gsn == 0 ? x1 : x2, xo = gsn == 0 ? x2 : x1;
    const float2 cb = mfp_combine(xg, mf_phasor((double)K * pf.thg), xo, mf_phasor((double)K * pf.tho),
                                  false, in_ov, pf.frg, pf.fro, gsn, w1, w2);
    macc += (double)norm_sqr(cb);
    if (cb_mid != nullptr) *cb_mid = cb;
}

// Consumer c of the phased kernel: the slots of producers c and c + 4,
// alternating, each slot's FPS frames in turn (as mf3_consumer and
// mfs_consumer).
template <int FPS>
__device__ void mfp_consumer(const SectorArgs &a, MfShared3 &sh, int c, int l, int64_t fa0,
                             int64_t fa1, int64_t fb0, int64_t fb1,
                             double (&acc)[2 * mfs::Shape<FPS>::NP], double &macc) {
    using S = mfs::Shape<FPS>;
    constexpr int M = S::MF, NP = S::NP, H = M + 1;
    const int gsn = a.grid;
    const int gs = a.a1s, ge = a.a1s + a.ov;   // a1s == a2s on this path
    uint32_t m_ov = 0;
#pragma unroll
    for (int s = 0; s <= 2 * NP; ++s) {
        const int k = mfs_bin<FPS>(s, l);
        if (k >= gs && k < ge) m_ov |= 1u << s;
    }
    const bool mid_ov = (m_ov >> (2 * NP)) & 1u;
    const float w1 = (float)a.w1, w2 = (float)a.w2;
    float2 *cb = ((a.flags & FLAG_RAW_OUT) != 0) ? a.cb_out : nullptr;
    const int64_t na = fa1 - fa0, nb = fb1 - fb0;
    const int64_t n = na > nb ? na : nb;
    const int q2 = c + mf::CONS;
    bool ok = true;
    for (int64_t g = 0; g < n; g += mf::FB) {
        const int cnt = (int)((n - g) < mf::FB ? (n - g) : mf::FB);
        MfPhFrame pa = {}, pb = {};
        if (l < cnt && g + l < na) pa = mfp_frame(a.first_frame + fa0 + g + l, a);
        if (l < cnt && g + l < nb) pb = mfp_frame(a.first_frame + fb0 + g + l, a);
        float2 ua1 = make_float2(0.f, 0.f), ua2 = ua1, ub1 = ua1, ub2 = ua1;   // Z'[M/2] of frame g + l
        const int ns = (cnt + FPS - 1) / FPS;
        for (int js = 0; js < ns; ++js) {
            const int i = (int)(g / FPS) + js;
            const int64_t fi = (int64_t)FPS * i;   // the slot's first frame (offset in each range)
            if (fi < na) {
                mf_wait_ge(&sh.prod[c], i + 1, ok, &sh.broken);
#pragma unroll
                for (int q = 0; q < FPS; ++q) {
                    if (fi + q < na) {
                        const int jf = js * FPS + q;
                        const float2 *z1 = &sh.z[c][0][q * M], *z2 = &sh.z[c][1][q * M];
                        mfp_mac<M, NP>(z1, z2, sh.twu, pa, jf, l, gsn, m_ov, w1, w2,
                                       cb ? cb + (fa0 + fi + q) * H : nullptr, acc);
                        const float2 u1 = z1[M / 2], u2 = z2[M / 2];
                        if (l == jf) {
                            ua1 = u1;
                            ua2 = u2;
                        }
                    }
                }
                mf_publish(&sh.cons[c], i + 1);
            }
            if (fi < nb) {
                mf_wait_ge(&sh.prod[q2], i + 1, ok, &sh.broken);
#pragma unroll
                for (int q = 0; q < FPS; ++q) {
                    if (fi + q < nb) {
                        const int jf = js * FPS + q;
                        const float2 *z1 = &sh.z[q2][0][q * M], *z2 = &sh.z[q2][1][q * M];
                        mfp_mac<M, NP>(z1, z2, sh.twu, pb, jf, l, gsn, m_ov, w1, w2,
                                       cb ? cb + (fb0 + fi + q) * H : nullptr, acc);
                        const float2 u1 = z1[M / 2], u2 = z2[M / 2];
                        if (l == jf) {
                            ub1 = u1;
                            ub2 = u2;
                        }
                    }
                }
                mf_publish(&sh.cons[q2], i + 1);
            }
        }
        if (l < cnt && g + l < na)
            mfp_mid(pa, ua1, ua2, gsn, mid_ov, w1, w2, cb ? cb + (fa0 + g + l) * H + M / 2 : nullptr, macc,
                    M / 2);
        if (l < cnt && g + l < nb)
            mfp_mid(pb, ub1, ub2, gsn, mid_ov, w1, w2, cb ? cb + (fb0 + g + l) * H + M / 2 : nullptr, macc,
                    M / 2);
    }
}

// accumulate = 0: the workgroup's partial slot is set; 1: its |cb|^2 sums
// are added (a sector in chunks that reuse the slots in order).
template <int FPS>
__global__ __launch_bounds__(mf::THREADS3) void k_phased_mf3(SectorArgs a, int accumulate) {
    __shared__ MfShared3 sh;
    using S = mfs::Shape<FPS>;
    constexpr int M = S::MF, NS = 2 * S::NP;
    constexpr int NT = mf::THREADS3;
    const int tid = threadIdx.x;
    const int l = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    mf3_setup(a, sh, tid);
    __syncthreads();
    const int64_t wg0 = (int64_t)blockIdx.x * a.frames_per_wg;
    int64_t wg1 = wg0 + a.frames_per_wg;
    if (wg1 > a.nf) wg1 = a.nf;
    const int64_t per = (a.frames_per_wg + mf::PROD - 1) / mf::PROD;
    auto range = [&](int p, int64_t &r0, int64_t &r1) {
        r0 = wg0 + per * p;
        if (r0 > wg1) r0 = wg1;
        r1 = (r0 + per < wg1) ? r0 + per : wg1;
    };
    double acc[NS], macc = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) acc[s] = 0.0;
    if (wv < mf::PROD) {
        int64_t f0, f1;
        range(wv, f0, f1);
        if constexpr (FPS == 1) mf3_producer(a, sh, wv, l, f0, f1);
        else mfs_producer<FPS>(a, sh, wv, l, f0, f1);
    } else {
        const int c = wv - mf::PROD;
        int64_t fa0, fa1, fb0, fb1;
        range(c, fa0, fa1);
        range(c + mf::CONS, fb0, fb1);
        mfp_consumer<FPS>(a, sh, c, l, fa0, fa1, fb0, fb1, acc, macc);
    }

    // ---- workgroup partial: consumers add their bins in a fixed order
    __syncthreads();
    double *red = reinterpret_cast<double *>(&sh.z[0][0][0]);   // [M + 1]
    double *mids = reinterpret_cast<double *>(&sh.z[4][0][0]);  // [CONS][64] lanes' bin-M/2 sums
    constexpr int H = M + 1;
    if (wv >= mf::PROD) mids[(wv - mf::PROD) * 64 + l] = macc;
    __syncthreads();
    for (int w = mf::PROD; w < mf::PROD + mf::CONS; ++w) {
        if (wv == w) {
            const bool first = w == mf::PROD;
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const int k = mfs_bin<FPS>(s, l);
                red[k] = (first ? 0.0 : red[k]) + acc[s];
            }
            if (l == 0) {
                const double *mc = mids + (w - mf::PROD) * 64;
                double sum = 0.0;
                for (int i = 0; i < 64; ++i) sum += mc[i];   // lanes in order
                red[M / 2] = (first ? 0.0 : red[M / 2]) + sum;
            }
        }
        __syncthreads();
    }
    double *out = a.part + (size_t)blockIdx.x * kNumProducts * H;
    const double bad = sh.broken ? __builtin_nan("") : 0.0;   // NaN marks a failed hand-over
    if (sh.broken && tid == 0 && a.err)
        __hip_atomic_fetch_or(a.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int k = tid; k < H; k += NT) {
        out[P_PH * H + k] = (accumulate ? out[P_PH * H + k] : 0.0) + red[k] + bad;
        if (!accumulate) {   // the other products stay zero over the later chunks
            out[P_11 * H + k] = 0.0;
            out[P_22 * H + k] = 0.0;
            out[P_12R * H + k] = 0.0;
            out[P_12I * H + k] = 0.0;
        }
    }
}

hipError_t launch_phased_mf3(const SectorArgs &a, int accumulate, int n_wg, hipStream_t s) {
    if ((a.N != mf::N && a.N != mf::N / 2 && a.N != mf::N / 4) || a.mf_tab == nullptr || a.a1s != a.a2s)
        return hipErrorInvalidValue;
    if (n_wg <= 0) return hipSuccess;
    if (a.N == mf::N)
        hipLaunchKernelGGL(k_phased_mf3<1>, dim3(n_wg), dim3(mf::THREADS3), 0, s, a, accumulate);
    else if (a.N == mf::N / 2)
        hipLaunchKernelGGL(k_phased_mf3<2>, dim3(n_wg), dim3(mf::THREADS3), 0, s, a, accumulate);
    else
        hipLaunchKernelGGL(k_phased_mf3<4>, dim3(n_wg), dim3(mf::THREADS3), 0, s, a, accumulate);
    return hipGetLastError();
}

// ------------------------------------------- N = 2048: k_sector_mf4
// The producer/consumer organization for M = 1024 = 32 x 32 (m = 32 m1 + m2,
// k = k1 + 32 k2).  Eight producers, each one station of one of four frame
// streams (producer p: stream p % 4, station p / 4), so that a producer's
// slot holds one station-frame (8 KB), as at N = 1024:
//   stage 1: the DFT-32 over m1 as FP16 MFMA GEMMs on two 16-row tiles
//            (m2 = 16 rt + l % 16), each B operand read feeding both tiles;
//   twiddle: W1024^{m2 k1}, stored transposed (column k1: the 32 values m2
//            in 256 bytes, 16-byte pairs XOR-swizzled by k1 & 7);
//   stage 2: lanes l and l + 32 share column k1 = l % 32; each reads the
//            whole column and forms its half h2 = l / 32 of the DFT-32 by
//            decimation in frequency (x[t] + x[t+16], or (x[t] - x[t+16])
//            W32^t), then that half's DFT-16 in registers, giving
//            Z/2[k1 + 32 (2 k' + h2)].
// Four consumers split the bins: consumer c takes the pairs (p, M - p),
// p = l + 64 j for j = 2c and 2c + 1, of every frame of every stream (32
// FP64 accumulator registers each); consumer 0 also bin M/2.  A slot is
// released when all four consumers have read it: cons[p] counts four per
// frame.
namespace mf4 {
constexpr int N = 2048, M = 1024, R2 = 32;
constexpr int STREAMS = 4, PROD = 8, CONS = 4;
constexpr int THREADS = 64 * (PROD + CONS);
constexpr int SROW = N + 2 * mf::PAD;
}  // namespace mf4

struct MfShared4 {
    uint4 b1[16 * 64];
    float tw2[32 * 64];
    float twu[16 * 64];
    uint32_t slut[2][4 * 256];
    uint16_t dlut[2][MFT_DLUT_HALVES];
    uint16_t stage[mf4::PROD][mf4::SROW];      // each producer's station row
    alignas(16) float2 z[mf4::PROD][mf4::M];   // one station-frame per producer
    float2 mid_u[mf4::STREAMS][mf::FB][2];     // consumer 0: Z'[M/2] per stream, batch frame, station
    int32_t prod[mf4::PROD], cons[mf4::PROD];
    int32_t broken;
};

__device__ __forceinline__ int mf4_tidx(int k1, int pair) { return k1 * 32 + 2 * (pair ^ (k1 & 7)); }

// W1024^{m2 k1} for rows m2 = 16 rt + 4g + r of column half h.
__device__ __forceinline__ void mf4_load_tw(const float *tw2, int rt, int h, int l, float2 tw[4]) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int idx = ((rt * 2 + h) * 4 + r) * 2;
        tw[r] = make_float2(tw2[idx * 64 + l], tw2[(idx + 1) * 64 + l]);
    }
}

__device__ __forceinline__ void mf4_twiddle_store(f4 re, f4 im, const float2 tw[4], float2 *slot,
                                                  int rt, int h, int l) {
    float2 v[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = c_mul(make_float2(re[r], im[r]), tw[r]);
    const int k1 = (l & 15) + 16 * h, pair = 2 * (l >> 4) + 8 * rt;   // m2 = 16 rt + 4g + r
    *reinterpret_cast<float4 *>(slot + mf4_tidx(k1, pair)) = make_float4(v[0].x, v[0].y, v[1].x, v[1].y);
    *reinterpret_cast<float4 *>(slot + mf4_tidx(k1, pair + 1)) = make_float4(v[2].x, v[2].y, v[3].x, v[3].y);
}

// cos(pi t / 16), t = 0..8.
__device__ constexpr float kCos16[9] = {1.0f, 0.98078528040323044913f, 0.92387953251128675613f,
                                        0.83146961230254523708f, 0.70710678118654752440f,
                                        0.55557023301960222474f, 0.38268343236508977173f,
                                        0.19509032201612826785f, 0.0f};

// x * W32^t, W32 = exp(-2 pi i / 32), t = 1..15.
template <int t>
__device__ __forceinline__ float2 mf_w32(float2 x) {
    if constexpr (t == 8) {
        return make_float2(x.y, -x.x);
    } else {
        constexpr float wr = t <= 8 ? kCos16[t] : -kCos16[16 - t];
        constexpr float wi = t <= 8 ? -kCos16[8 - t] : -kCos16[t - 8];
        return make_float2(x.x * wr - x.y * wi, x.x * wi + x.y * wr);
    }
}

template <int t>
__device__ __forceinline__ void mf_w32_step(float2 (&d)[16]) {
    if constexpr (t < 16) {
        d[t] = mf_w32<t>(d[t]);
        mf_w32_step<t + 1>(d);
    }
}

// Stage 2 of one station-frame (M = 1024) from the transposed slot.
__device__ __forceinline__ void mf4_columns(float2 *slot, int l) {
    const int k1 = l & 31, h2 = l >> 5;
    const float sgn = h2 ? -1.0f : 1.0f;
    float2 d[16];
#pragma unroll
    for (int j = 0; j < 8; ++j) {   // pairs j (m2 = 2j, 2j + 1) and j + 8 (m2 + 16)
        const float4 qa = *reinterpret_cast<const float4 *>(slot + mf4_tidx(k1, j));
        const float4 qb = *reinterpret_cast<const float4 *>(slot + mf4_tidx(k1, j + 8));
        d[2 * j] = make_float2(fmaf(sgn, qb.x, qa.x), fmaf(sgn, qb.y, qa.y));
        d[2 * j + 1] = make_float2(fmaf(sgn, qb.z, qa.z), fmaf(sgn, qb.w, qa.w));
    }
    wave_lds_sync();   // every lane's column is read before Z overwrites the slot
    if (h2) mf_w32_step<1>(d);
    mf_dft16_out(d, slot + k1 + 32 * h2, 64);
}

__device__ void mf4_producer(const SectorArgs &a, MfShared4 &sh, int p, int l, int64_t f0, int64_t f1) {
    const int st = p / mf4::STREAMS;
    const int bits = a.st[st].bits, kind = a.st[st].shuffle_kind, nw = a.st[st].nwords;
    const int64_t bpf = a.st[st].bpf;
    const uint8_t *raw = st ? a.raw2 : a.raw1;
    uint16_t *row = sh.stage[p];
    float2 *slot = sh.z[p];
    uint32_t cur[4] = {0, 0, 0, 0};
    bool ok = true;
    if (f0 < f1) {
        mf_load(raw + f0 * bpf, nw, l, cur);
        mf_stage(cur, bits, kind, nw, sh.slut[st], sh.dlut[st], row + mf::PAD, l);
        if (f0 + 1 < f1) mf_load(raw + (f0 + 1) * bpf, nw, l, cur);
        wave_lds_sync();
    }
    for (int64_t fb = f0; fb < f1; fb += mf::FB) {
        const int nb = (int)((f1 - fb) < mf::FB ? (f1 - fb) : mf::FB);
        MfFrame pf = {};
        if (l < nb) pf = mf_frame(a.first_frame + fb + l, a);
        for (int jf = 0; jf < nb; ++jf) {
            const int64_t f = fb + jf;
            const int32_t i = (int32_t)(f - f0);
            uint32_t nxt[4] = {0, 0, 0, 0};
            if (f + 2 < f1) mf_load(raw + (f + 2) * bpf, nw, l, nxt);
            const int sh_int = __builtin_amdgcn_readlane(st ? pf.s2 : pf.s1, jf);
            half8 a00, a01, a10, a11;   // row tiles 0 and 1
            mf_build_a_t<mf4::R2, mf4::SROW>(row, sh_int, l, 0, a00, a01);
            mf_build_a_t<mf4::R2, mf4::SROW>(row, sh_int, l, 1, a10, a11);
#pragma unroll 1
            for (int h = 0; h < 2; ++h) {
                f4 y0[2], y1[2];   // tiles h (re) and 2 + h (im) of row tiles 0 and 1
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const int t = h + 2 * q;
                    const half8 bh0 = ld_b1(sh.b1, t * 2 + 0, l), bl0 = ld_b1(sh.b1, t * 2 + 1, l);
                    const half8 bh1 = ld_b1(sh.b1, (4 + t) * 2 + 0, l), bl1 = ld_b1(sh.b1, (4 + t) * 2 + 1, l);
                    y0[q] = mf_stage1(a00, a01, bh0, bl0, bh1, bl1);
                    y1[q] = mf_stage1(a10, a11, bh0, bl0, bh1, bl1);
                }
                float2 tw0[4], tw1[4];
                mf4_load_tw(sh.tw2, 0, h, l, tw0);
                mf4_load_tw(sh.tw2, 1, h, l, tw1);
                // the slot was last read by the four consumers for frame i - 1
                if (h == 0) mf_wait_ge(&sh.cons[p], mf4::CONS * i, ok, &sh.broken);
                mf4_twiddle_store(y0[0], y0[1], tw0, slot, 0, h, l);
                mf4_twiddle_store(y1[0], y1[1], tw1, slot, 1, h, l);
            }
            wave_lds_sync();
            mf4_columns(slot, l);
            mf_publish(&sh.prod[p], i + 1);
            if (f + 1 < f1) {   // decode frame f+1 (the staging row is this wave's own)
                wave_lds_sync();
                mf_stage(cur, bits, kind, nw, sh.slut[st], sh.dlut[st], row + mf::PAD, l);
            }
            wave_lds_sync();
#pragma unroll
            for (int u = 0; u < 4; ++u) cur[u] = nxt[u];
        }
    }
}

// Consumer c's two bin pairs of one frame (slots z1, z2: the two stations).
// Local slot s: j = 2c + s/2, p = l + 64 j, bin p (s even) or M - p (odd).
template <bool FULL>
__device__ __forceinline__ void mf4_mac(const float2 *z1, const float2 *z2, const float *twu,
                                        const MfFrame &pf, int jf, int c, int l, bool do_acf,
                                        bool do_xcf, uint32_t m11, uint32_t m22, uint32_t m_ov,
                                        float k11, float k22, float kM, double (&acc)[4][4]) {
    constexpr int M = mf4::M;
    const double phw = readlane_d(pf.phw, jf);
    const double dth = readlane_d(pf.dth, jf);
    const float2 F1 = make_float2(readlane_f(pf.F1.x, jf), readlane_f(pf.F1.y, jf));
    const float2 Cm = make_float2(readlane_f(pf.C.x, jf), readlane_f(pf.C.y, jf));
    const float r0 = (float)wrap_turns(fma((double)(l + 128 * c), dth, phw));
    float2 P = make_float2(__builtin_amdgcn_cosf(r0), __builtin_amdgcn_sinf(r0));
#pragma unroll
    for (int jj = 0; jj < 2; ++jj) {
        const int j = 2 * c + jj;
        const int p = l + 64 * j;
        const int pm = (M - p) & (M - 1);
        const float2 w = make_float2(twu[(2 * j) * 64 + l], twu[(2 * j + 1) * 64 + l]);
        float2 x1p, x1m, x2p, x2m;
        untangle_pair(z1[p], z1[pm], w, x1p, x1m);
        untangle_pair(z2[p], z2[pm], w, x2p, x2m);
        const int sp = 2 * jj, sm = 2 * jj + 1;
        const bool j0 = jj == 0 && c == 0;   // lane 0's slot 1 is bin M
        if (do_acf) {
            if (FULL) {
                acc[sp][0] += (double)norm_sqr(x1p);
                acc[sp][1] += (double)norm_sqr(x2p);
                acc[sm][0] += (double)(j0 ? k11 * norm_sqr(x1m) : norm_sqr(x1m));
                acc[sm][1] += (double)(j0 ? k22 * norm_sqr(x2m) : norm_sqr(x2m));
            } else {
                acc[sp][0] += (double)(((m11 >> sp) & 1u) ? norm_sqr(x1p) : 0.0f);
                acc[sp][1] += (double)(((m22 >> sp) & 1u) ? norm_sqr(x2p) : 0.0f);
                acc[sm][0] += (double)(((m11 >> sm) & 1u) ? norm_sqr(x1m) : 0.0f);
                acc[sm][1] += (double)(((m22 >> sm) & 1u) ? norm_sqr(x2m) : 0.0f);
            }
        }
        if (do_xcf) {
            float2 Pp, Pm;
            if (FULL) {
                Pp = P;
                Pm = c_mul(Cm, c_conj(P));
                if (j0) Pm = make_float2(kM * Pm.x, kM * Pm.y);
            } else {
                Pp = ((m_ov >> sp) & 1u) ? P : make_float2(0.f, 0.f);
                Pm = ((m_ov >> sm) & 1u) ? c_mul(Cm, c_conj(P)) : make_float2(0.f, 0.f);
            }
            const float2 vp = c_mul(c_mul(x1p, c_conj(x2p)), Pp);
            const float2 vm = c_mul(c_mul(x1m, c_conj(x2m)), Pm);
            acc[sp][2] += (double)vp.x;
            acc[sp][3] += (double)vp.y;
            acc[sm][2] += (double)vm.x;
            acc[sm][3] += (double)vm.y;
        }
        P = c_mul(P, F1);
    }
}

__device__ __forceinline__ int mf4_bin(int c, int s, int l) {
    const int p = l + 64 * (2 * c + (s >> 1));
    return (s & 1) ? mf4::M - p : p;
}

// One lane of the wave adds 1 to *flag (release: this wave's reads of the
// slot are complete).
__device__ __forceinline__ void mf_release_add(int32_t *flag, int l) {
    if (l == 0) __hip_atomic_fetch_add(flag, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <bool FULL>
__device__ void mf4_consumer(const SectorArgs &a, MfShared4 &sh, int c, int l, const int64_t (&fs0)[4],
                             const int64_t (&fs1)[4], double (&acc)[4][4], double (&macc)[4]) {
    const bool do_acf = (a.flags & FLAG_ACF) != 0;
    const bool do_xcf = (a.flags & FLAG_XCF) != 0;
    const bool ov_only = (a.flags & FLAG_ACF_OVERLAP) != 0;
    const int gs = a.a1s, ge = a.a1s + a.ov;   // a1s == a2s on this path
    uint32_t m_grid = 0, m_ov = 0;
#pragma unroll
    for (int s = 0; s < 5; ++s) {   // local slots 0..3, bit 4: bin M/2
        const int k = s == 4 ? mf4::M / 2 : mf4_bin(c, s, l);
        const bool in_ov = k >= gs && k < ge;
        if (!ov_only || in_ov) m_grid |= 1u << s;
        if (in_ov) m_ov |= 1u << s;
    }
    const uint32_t m11 = a.grid == 0 ? m_grid : m_ov;
    const uint32_t m22 = a.grid == 0 ? m_ov : m_grid;
    const float kM = (l == 0) ? 0.0f : 1.0f;   // FULL: bin M (lane 0, consumer 0, slot 1)
    const float k11 = a.grid == 0 ? 1.0f : kM, k22 = a.grid == 0 ? kM : 1.0f;
    const bool in11 = (m11 >> 4) & 1u, in22 = (m22 >> 4) & 1u, in_ovm = (m_ov >> 4) & 1u;
    int64_t len[4], n = 0;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        len[s] = fs1[s] - fs0[s];
        n = len[s] > n ? len[s] : n;
    }
    bool ok = true;
    for (int64_t g = 0; g < n; g += mf::FB) {
        const int cnt = (int)((n - g) < mf::FB ? (n - g) : mf::FB);
        MfFrame pf[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            pf[s] = MfFrame{};
            if (l < cnt && g + l < len[s]) pf[s] = mf_frame(a.first_frame + fs0[s] + g + l, a);
        }
        for (int jf = 0; jf < cnt; ++jf) {
            const int32_t i = (int32_t)(g + jf);
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                if (i >= len[s]) continue;
                mf_wait_ge(&sh.prod[s], i + 1, ok, &sh.broken);
                mf_wait_ge(&sh.prod[s + 4], i + 1, ok, &sh.broken);
                mf4_mac<FULL>(sh.z[s], sh.z[s + 4], sh.twu, pf[s], jf, c, l, do_acf, do_xcf, m11, m22,
                              m_ov, k11, k22, kM, acc);
                if (c == 0 && l == 0) {   // bin M/2 of this frame, formed after the batch
                    sh.mid_u[s][jf][0] = sh.z[s][mf4::M / 2];
                    sh.mid_u[s][jf][1] = sh.z[s + 4][mf4::M / 2];
                }
                mf_release_add(&sh.cons[s], l);       // this wave's reads of both slots are done
                mf_release_add(&sh.cons[s + 4], l);
            }
        }
        if (c == 0) {
            wave_lds_sync();   // lane 0's stores of the batch before the lanes' loads
#pragma unroll
            for (int s = 0; s < 4; ++s)
                if (l < cnt && g + l < len[s])
                    mf_mid(pf[s], sh.mid_u[s][l][0], sh.mid_u[s][l][1], do_acf, do_xcf, FULL || in11,
                           FULL || in22, FULL || in_ovm, macc, mf4::M / 2);
            wave_lds_sync();   // and those loads before the next batch's stores
        }
    }
}

__global__ __launch_bounds__(mf4::THREADS) void k_sector_mf4(SectorArgs a) {
    __shared__ MfShared4 sh;
    constexpr int M = mf4::M;
    constexpr int NT = mf4::THREADS;
    const int tid = threadIdx.x;
    const int l = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint8_t *tab = reinterpret_cast<const uint8_t *>(a.mf_tab);
    {
        const uint4 *g = reinterpret_cast<const uint4 *>(tab + MFT4_B1);
        for (int i = tid; i < 16 * 64; i += NT) sh.b1[i] = g[i];
        const uint32_t *gs = reinterpret_cast<const uint32_t *>(tab + MFT4_SLUT);
        for (int i = tid; i < 2 * 4 * 256; i += NT) (&sh.slut[0][0])[i] = gs[i];
        const uint32_t *gd = reinterpret_cast<const uint32_t *>(tab + MFT4_DLUT);
        for (int i = tid; i < 2 * MFT_DLUT_HALVES / 2; i += NT)
            reinterpret_cast<uint32_t *>(&sh.dlut[0][0])[i] = gd[i];
        uint32_t *st32 = reinterpret_cast<uint32_t *>(&sh.stage[0][0]);
        for (int i = tid; i < mf4::PROD * mf4::SROW / 2; i += NT) st32[i] = 0u;
        const float *gt = reinterpret_cast<const float *>(tab + MFT4_TW2);
        const float *gu = reinterpret_cast<const float *>(tab + MFT4_TWU);
        for (int i = tid; i < 32 * 64; i += NT) sh.tw2[i] = gt[i];
        for (int i = tid; i < 16 * 64; i += NT) sh.twu[i] = gu[i];
        if (tid < mf4::PROD) {
            sh.prod[tid] = 0;
            sh.cons[tid] = 0;
        }
        if (tid == 0) sh.broken = 0;
    }
    __syncthreads();

    // frames of stream s: a contiguous quarter of the workgroup's range
    const int64_t wg0 = (int64_t)blockIdx.x * a.frames_per_wg;
    int64_t wg1 = wg0 + a.frames_per_wg;
    if (wg1 > a.nf) wg1 = a.nf;
    const int64_t per = (a.frames_per_wg + mf4::STREAMS - 1) / mf4::STREAMS;
    auto range = [&](int s, int64_t &r0, int64_t &r1) {
        r0 = wg0 + per * s;
        if (r0 > wg1) r0 = wg1;
        r1 = (r0 + per < wg1) ? r0 + per : wg1;
    };

    double acc[4][4], macc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[s][c] = 0.0;
    if (wv < mf4::PROD) {
        int64_t f0, f1;
        range(wv % mf4::STREAMS, f0, f1);
        mf4_producer(a, sh, wv, l, f0, f1);
    } else {
        const int c = wv - mf4::PROD;
        int64_t fs0[4], fs1[4];
        for (int s = 0; s < 4; ++s) range(s, fs0[s], fs1[s]);
        const bool full = (a.flags & FLAG_ACF_OVERLAP) == 0 && a.a1s == 0 && a.ov == M;
        if (full) mf4_consumer<true>(a, sh, c, l, fs0, fs1, acc, macc);
        else mf4_consumer<false>(a, sh, c, l, fs0, fs1, acc, macc);
    }

    // ---- workgroup partial: every bin belongs to one consumer lane
    __syncthreads();
    constexpr int H = M + 1;
    double *red = reinterpret_cast<double *>(&sh.z[0][0]);    // [4][H]
    double *mids = reinterpret_cast<double *>(&sh.z[5][0]);   // [4][64] consumer 0's lane sums
    if (wv >= mf4::PROD) {
        const int c = wv - mf4::PROD;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int k = mf4_bin(c, s, l);
#pragma unroll
            for (int q = 0; q < 4; ++q) red[q * H + k] = acc[s][q];
        }
        if (c == 0)
#pragma unroll
            for (int q = 0; q < 4; ++q) mids[q * 64 + l] = macc[q];
    }
    __syncthreads();
    if (tid == 0)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            double sum = 0.0;
            for (int i = 0; i < 64; ++i) sum += mids[q * 64 + i];   // lanes in order
            red[q * H + M / 2] = sum;
        }
    __syncthreads();
    double *out = a.part + (size_t)blockIdx.x * kNumProducts * H;
    const double bad = sh.broken ? __builtin_nan("") : 0.0;   // NaN marks a failed hand-over
    if (sh.broken && tid == 0 && a.err)   // and the plan's error word reports it to the host
        __hip_atomic_fetch_or(a.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int k = tid; k < H; k += NT) {
        out[P_PH * H + k] = bad;
        out[P_11 * H + k] = red[0 * H + k] + bad;
        out[P_22 * H + k] = red[1 * H + k] + bad;
        out[P_12R * H + k] = red[2 * H + k] + bad;
        out[P_12I * H + k] = red[3 * H + k] + bad;
    }
}

hipError_t launch_sector_mfma(const SectorArgs &a, int precision, int n_wg, hipStream_t s) {
    if ((a.N != mf::N && a.N != mf4::N && a.N != mf::N / 4 && a.N != mf::N / 2) || a.mf_tab == nullptr ||
        precision != 1 || a.a1s != a.a2s)
        return hipErrorInvalidValue;
    if (a.N == mf::N / 4 || a.N == mf::N / 2) {   // packed frames, one organization
        if (a.N == mf::N / 4) hipLaunchKernelGGL(k_sector_mfs<4>, dim3(n_wg), dim3(mf::THREADS3), 0, s, a);
        else hipLaunchKernelGGL(k_sector_mfs<2>, dim3(n_wg), dim3(mf::THREADS3), 0, s, a);
        return hipGetLastError();
    }
    if (a.N == mf4::N) {   // one organization at N = 2048
        hipLaunchKernelGGL(k_sector_mf4, dim3(n_wg), dim3(mf4::THREADS), 0, s, a);
        return hipGetLastError();
    }
    // variant 1: producer/consumer waves; 2: the same with FP32 batch sums
    // in the consumers (needs a.cons_acc); 0: whole-frame waves
    if (a.variant == 1) hipLaunchKernelGGL(k_sector_mf3<false>, dim3(n_wg), dim3(mf::THREADS3), 0, s, a);
    else if (a.variant == 2) {
        if (a.cons_acc == nullptr) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_sector_mf3<true>, dim3(n_wg), dim3(mf::THREADS3), 0, s, a);
    } else hipLaunchKernelGGL(k_sector_mf2, dim3(n_wg), dim3(512), 0, s, a);
    return hipGetLastError();
}

int mfma_wg_threads() { return 512; }

// Forward R2C of frames through the same two-stage MFMA FFT (test entry):
// one wave per frame.  Inputs must be exact in FP16.
__global__ __launch_bounds__(64) void k_rfft_mfma(SectorArgs a, const float *in, int64_t nframes,
                                                  float2 *out) {
    // variant 1: two frames per workgroup through the column DFT-16 (the
    // producer/consumer kernel's stage 2, the frames in the station slots)
    __shared__ uint4 b1[16 * 64];
    __shared__ uint16_t row[2][mf::SROW];
    __shared__ alignas(16) float2 z[2][mf::M];
    const int l = threadIdx.x;
    const int per = a.variant == 1 ? 2 : 1;
    const int64_t f0 = (int64_t)blockIdx.x * per;
    const uint8_t *tab = reinterpret_cast<const uint8_t *>(a.mf_tab);
    const uint4 *g = reinterpret_cast<const uint4 *>(tab + MFT_B1);
    for (int i = l; i < 16 * 64; i += 64) b1[i] = g[i];
    for (int s = 0; s < per; ++s)
        for (int i = l; i < mf::SROW; i += 64) {
            const int n = i - mf::PAD;
            const bool in_frame = n >= 0 && n < mf::N && f0 + s < nframes;
            const float v = in_frame ? in[(f0 + s) * mf::N + n] : 0.0f;
            row[s][i] = __builtin_bit_cast(uint16_t, (_Float16)v);
        }
    const float *ga = reinterpret_cast<const float *>(tab + MFT_A2);
    const float *gt = reinterpret_cast<const float *>(tab + MFT_TW2);
    const float *gu = reinterpret_cast<const float *>(tab + MFT_TWU);
    __syncthreads();
    for (int s = 0; s < per; ++s) {
        half8 a0, a1;
        mf_build_a(row[s], 0, l, a0, a1);
        f4 y[4];
#pragma unroll
        for (int t = 0; t < 4; ++t)
            y[t] = mf_stage1(a0, a1, ld_b1(b1, t * 2, l), ld_b1(b1, t * 2 + 1, l),
                             ld_b1(b1, (4 + t) * 2, l), ld_b1(b1, (4 + t) * 2 + 1, l));
        if (a.variant == 1) mf_twiddle_store(y, gt, &z[0][0], s, l);
        else mf_stage2(y, gt, ga, z[0], l);
    }
    if (a.variant == 1) {
        wave_lds_sync();
        mf_columns_dft16(&z[0][0], l);
    }
    __syncthreads();
    for (int s = 0; s < per; ++s) {
        if (f0 + s >= nframes) break;
        float2 *o = out + (f0 + s) * (mf::M + 1);
        const float2 *zs = z[s];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int p = l + 64 * j;
            float2 xp, xm;
            untangle_pair(zs[p], zs[(mf::M - p) & (mf::M - 1)],
                          make_float2(gu[(2 * j) * 64 + l], gu[(2 * j + 1) * 64 + l]), xp, xm);
            o[p] = xp;
            o[mf::M - p] = xm;
        }
        if (l == 0) o[mf::M / 2] = make_float2(2.0f * zs[mf::M / 2].x, -2.0f * zs[mf::M / 2].y);
    }
}

hipError_t launch_rfft_mfma(const SectorArgs &a, const float *in, int64_t nframes, float2 *out,
                            hipStream_t s) {
    if (a.N != mf::N || a.mf_tab == nullptr) return hipErrorInvalidValue;
    if (nframes <= 0) return hipSuccess;
    const int64_t per = a.variant == 1 ? 2 : 1;
    hipLaunchKernelGGL(k_rfft_mfma, dim3((unsigned)((nframes + per - 1) / per)), dim3(64), 0, s, a, in,
                       nframes, out);
    return hipGetLastError();
}

// ------------------------------- N >= 32768: first pass of the HBM four-step
// The column DFTs of length L = 32 R of the first pass (m = M2 m1 + m2, one
// column per m2, L = M1), split m1 = v + R u, k1 = k' + 32 k'':
//   A_v[k'] = sum_u W32^{u k'} x[v + R u]     FP16 hi + lo MFMA GEMM: rows =
//                                              16 columns at one v, K = u
//   Y[k' + 32 k''] = sum_v W_R^{v k''} (W_L^{v k'} A_v[k'])      FP32
// and Y[k1] W_M^{m2 k1} stored at out[m2 + M2 k1], as the FP32 pass does.
// Wave v forms A_v of the workgroup's 16 columns; the DFT-R runs after an
// LDS exchange, one (column, k') pair per thread.  The inputs are decoded
// levels, exact in FP16 (the plan checks this): FP32 pairs converted exactly
// (IN = float2), or FP16 pairs written so by the decode (IN = uint32_t).
template <int R, typename IN>
__global__ __launch_bounds__(64 * R) void k_fft_pass1_mfma(const IN *in, float2 *out,
                                                           int64_t nbatch, int M2,
                                                           const uint4 *b1g, const float2 *tw_m,
                                                           int M) {
    constexpr int KP = 33;   // padded k' row: lanes of consecutive columns hit distinct banks
    __shared__ uint4 b1[16 * 64];
    __shared__ float2 sb[R][16][KP];
    const int tid = threadIdx.x;
    const int l = tid & 63;
    const int v = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int64_t blocks_per_frame = M2 / 16;
    const int64_t f = blockIdx.x / blocks_per_frame;
    const int c0 = (int)(blockIdx.x % blocks_per_frame) * 16;
    if (f >= nbatch) return;   // uniform per workgroup
    const IN *src = in + f * (int64_t)M;
    float2 *dst = out + f * (int64_t)M;
    for (int i = tid; i < 16 * 64; i += 64 * R) b1[i] = b1g[i];
    // A operand of K-step kap, lane l: row r = l % 16 (column c0 + r),
    // slots (re, im) of u = 16 kap + g + 4 q, g = l / 16, q = 0..3
    const int r = l & 15, g = l >> 4;
    uint32_t d[8];
#pragma unroll
    for (int kap = 0; kap < 2; ++kap)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int u = 16 * kap + g + 4 * q;
            const IN x = src[(c0 + r) + (int64_t)M2 * (v + R * u)];
            if constexpr (sizeof(IN) == 4)
                d[kap * 4 + q] = x;
            else
                d[kap * 4 + q] = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(x.x, x.y));
        }
    const half8 a0 = __builtin_bit_cast(half8, make_uint4(d[0], d[1], d[2], d[3]));
    const half8 a1 = __builtin_bit_cast(half8, make_uint4(d[4], d[5], d[6], d[7]));
    __syncthreads();   // B table in LDS
    f4 y[4];           // tile t: k' = l % 16 + 16 (t & 1), re (t < 2) | im; rows 4 g + i
#pragma unroll
    for (int t = 0; t < 4; ++t)
        y[t] = mf_stage1(a0, a1, ld_b1(b1, t * 2 + 0, l), ld_b1(b1, t * 2 + 1, l),
                         ld_b1(b1, (4 + t) * 2 + 0, l), ld_b1(b1, (4 + t) * 2 + 1, l));
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int i = 0; i < 4; ++i) sb[v][4 * g + i][r + 16 * h] = make_float2(y[h][i], y[2 + h][i]);
    __syncthreads();
    for (int pi = tid; pi < 16 * 32; pi += 64 * R) {
        const int rr = pi & 15, kp = pi >> 4;
        float2 a[R];
#pragma unroll
        for (int vv = 0; vv < R; ++vv) {
            a[vv] = sb[vv][rr][kp];
            if (vv > 0) a[vv] = c_mul(a[vv], tw_m[vv * kp * M2]);   // W_L^{v k'} = W_M^{v k' M2}
        }
        dftR<R>(a);
        const int64_t m2 = c0 + rr;
#pragma unroll
        for (int k2 = 0; k2 < R; ++k2) {
            const int64_t k1 = kp + 32 * k2;
            dst[m2 + k1 * M2] = c_mul(a[k2], tw_m[(m2 * k1) & (M - 1)]);   // M: a power of two
        }
    }
}

// One element (two consecutive samples, the first at src0 = 2m - shift of
// the shifted frame) from words staged per four-step row: lw is the local
// index of the word holding bit src0 * bits and off its bit offset.  The
// code width is fixed at compile time for 1, 2, 4 and 8 bits (0: the
// station's run-time width, at most 16 so that both codes fit one funnel
// shift).  Branch-free: the LSB flip is a sign mask per sample parity (fm0
// for the first sample, fm1 for the second; the parity of src0 is the
// shift's), and samples outside the frame are zero.
template <int BITS>
__device__ __forceinline__ float2 decode_pair_local(const uint32_t *words, const StationDev &st,
                                                    const float *lev_lds, int src0, int lw,
                                                    uint32_t off, int N, uint32_t fm0,
                                                    uint32_t fm1) {
    const uint32_t v = __builtin_amdgcn_alignbit(words[lw + 1], words[lw], off);
    float x0, x1;
    if constexpr (BITS == 1) {
        x0 = (v & 1u) ? st.lev4[1] : st.lev4[0];
        x1 = (v & 2u) ? st.lev4[1] : st.lev4[0];
    } else if constexpr (BITS == 2 || BITS == 4 || BITS == 8) {
        constexpr uint32_t mask = (1u << BITS) - 1u;
        x0 = lev_lds[v & mask];
        x1 = lev_lds[(v >> BITS) & mask];
    } else {
        x0 = level_of(v & st.code_mask, st, lev_lds);
        x1 = level_of((v >> st.bits) & st.code_mask, st, lev_lds);
    }
    x0 = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, x0) ^ fm0);
    x1 = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, x1) ^ fm1);
    return make_float2((unsigned)src0 < (unsigned)N ? x0 : 0.0f,
                       (unsigned)(src0 + 1) < (unsigned)N ? x1 : 0.0f);
}

// Word gw of a frame (little-endian, shuffle applied); zero outside it.
__device__ __forceinline__ uint32_t frame_word(const uint8_t *frame, const StationDev &st, int gw) {
    if (gw < 0 || gw >= st.nwords) return 0u;
    uint32_t v = 0;
    if (st.aligned) {
        v = reinterpret_cast<const uint32_t *>(frame)[gw];
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int b = 4 * gw + j;
            if (b < st.bpf) v |= (uint32_t)frame[b] << (8 * j);
        }
    }
    return shuffle_word(v, st);
}

// The same first pass with the frame decode fused in (chosen form, sector
// path, M1 = 128 = 32 x 4).  Workgroup (frame, part p) takes the columns
// m2 in [64 p, 64 p + 64) of all 128 rows m1 (element m = m2 + M2 m1): it
// stages in LDS, per row, the K = 4 bits + 1 packed words that hold those
// columns' samples (shuffle applied), and its four waves take one 16-column
// block each, with no workgroup barrier after the stage:
//  - for each v = 0..3 the wave decodes its A_v operand from the staged
//    words (levels, integer shift with zero fill, LSB flip) and runs the
//    DFT-32 GEMM with the operands in swapped roles (table x data), so lane
//    (r = l % 16, g = l / 16) receives column c0 + r at k' = 4 g + i (+ 16
//    for the second k' tile) for every v;
//  - the DFT-4 over v is then lane-local (no LDS exchange), and each store
//    instruction writes 16 consecutive columns of 4 rows k1 = k' + 32 k2.
// Y[k1][m2] W_M^{m2 k1} lands at out[m2 + M2 k1] as in k_fft_pass1_mfma; the
// output twiddle is W_M^{m2 k'} W_M^{32 m2 k2} (two table values).
// Row m1's words start at word m1 rw + base0 of the frame, rw = M2 bits / 16
// (the row's 2 M2 samples), base0 = floor((128 p - shift) bits / 32).
__device__ __forceinline__ f4 mf_stage1_t(half8 a0, half8 a1, half8 bh0, half8 bl0, half8 bh1,
                                          half8 bl1) {
    f4 d = {0.f, 0.f, 0.f, 0.f};
    d = __builtin_amdgcn_mfma_f32_16x16x32_f16(bl0, a0, d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_16x16x32_f16(bl1, a1, d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_16x16x32_f16(bh0, a0, d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_16x16x32_f16(bh1, a1, d, 0, 0, 0);
    return d;
}

template <int BITS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_large_pass1_fused(
    SectorArgs a, int station, const int64_t *shifts, float2 *out, int M2, int parts,
    const float2 *tw_m) {
    constexpr int R = 4;
    extern __shared__ uint32_t words[];   // 128 rows x K words, one zero word after them
    __shared__ uint4 b1[16 * 64];
    __shared__ float2 twv[R][32];         // W_128^{v k'}
    __shared__ float lev[256];
    const StationDev &st = a.st[station];
    const int bits = BITS ? BITS : st.bits;
    const int K = 4 * bits + 1, rw = M2 * bits / 16;
    const int M = a.M;
    const int tid = threadIdx.x;
    const int l = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    // workgroups are dealt to the 8 XCDs in turn: give each XCD a contiguous
    // run of (frame, part) so a frame's parts share one L2 (their rows'
    // words share cache lines)
    const int nb = (int)gridDim.x, xcd = (int)blockIdx.x & 7, k = (int)blockIdx.x >> 3;
    const int lb = xcd * (nb >> 3) + min(xcd, nb & 7) + k;
    const int64_t f = lb / parts;
    const int part = lb % parts;
    if (f >= a.nf) return;   // uniform per workgroup
    const uint4 *b1g = reinterpret_cast<const uint4 *>(a.mf_tab);
    for (int i = tid; i < 16 * 64; i += 256) b1[i] = b1g[i];
    if (tid < R * 32) twv[tid >> 5][tid & 31] = tw_m[(tid >> 5) * (tid & 31) * M2];
    if (st.bits >= 2 && st.bits <= 8)
        for (int i = tid; i < 256; i += 256) lev[i] = (i <= (int)st.code_mask) ? st.levels[i] : 0.0f;
    const int sh = clamp_shift(shifts[f * 2 + station], a.N);
    const int base0 = ((128 * part - sh) * bits) >> 5;   // arithmetic shift: floor
    {
        const uint8_t *frame = (station == 0 ? a.raw1 : a.raw2) + f * st.bpf;
        const int nst = 128 * K;
        for (int i0 = 0; i0 < nst; i0 += 256 * 8) {   // eight loads in flight per thread
            uint32_t v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int idx = i0 + tid + 256 * u;
                const int m1 = idx / K;
                v[u] = idx < nst ? frame_word(frame, st, m1 * rw + base0 + (idx - m1 * K)) : 0u;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int idx = i0 + tid + 256 * u;
                if (idx < nst) words[idx] = v[u];
            }
        }
        if (tid == 0) words[nst] = 0u;
    }
    __syncthreads();
    // LSB flip (odd samples of the unshifted frame): sample 2m - sh is odd iff sh is
    const uint32_t fm0 = (st.lsb && (sh & 1)) ? 0x80000000u : 0u;
    const uint32_t fm1 = (st.lsb && !(sh & 1)) ? 0x80000000u : 0u;
    const int r = l & 15, g = l >> 4;
    float2 *dst = out + f * (int64_t)M;
    {
        const int cb = part * 4 + w;
        const int c0 = cb * 16;
        const int m2 = c0 + r;
        // output twiddles: W_M^{m2 k'} for this lane's 8 k', W_M^{32 m2 k2}
        float2 wk[2][4], w32[R];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int i = 0; i < 4; ++i) wk[h][i] = tw_m[(m2 * (4 * g + i + 16 * h)) & (M - 1)];
#pragma unroll
        for (int k2 = 1; k2 < R; ++k2) w32[k2] = tw_m[(32 * m2 * k2) & (M - 1)];
        // the A_v operands of all four v (decoded once), then per k' half h
        // the two tiles (re, im) of every v, their DFT-4 over v and stores
        half8 av0[R], av1[R];
#pragma unroll
        for (int v = 0; v < R; ++v) {
            uint32_t d[8];
#pragma unroll
            for (int kap = 0; kap < 2; ++kap)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int u = 16 * kap + g + 4 * q;
                    const int m1 = v + R * u;
                    const int src0 = 2 * (m2 + M2 * m1) - sh;
                    const int bp = src0 * bits;
                    const float2 x = decode_pair_local<BITS>(words, st, lev, src0,
                                                             (bp >> 5) - base0 + m1 * (K - rw),
                                                             (uint32_t)bp & 31u, a.N, fm0, fm1);
                    d[kap * 4 + q] = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(x.x, x.y));
                }
            av0[v] = __builtin_bit_cast(half8, make_uint4(d[0], d[1], d[2], d[3]));
            av1[v] = __builtin_bit_cast(half8, make_uint4(d[4], d[5], d[6], d[7]));
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            f4 yr[R], yi[R];
#pragma unroll
            for (int v = 0; v < R; ++v) {
                yr[v] = mf_stage1_t(av0[v], av1[v], ld_b1(b1, h * 2 + 0, l), ld_b1(b1, h * 2 + 1, l),
                                    ld_b1(b1, (4 + h) * 2 + 0, l), ld_b1(b1, (4 + h) * 2 + 1, l));
                yi[v] = mf_stage1_t(av0[v], av1[v], ld_b1(b1, (2 + h) * 2 + 0, l),
                                    ld_b1(b1, (2 + h) * 2 + 1, l), ld_b1(b1, (6 + h) * 2 + 0, l),
                                    ld_b1(b1, (6 + h) * 2 + 1, l));
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int kp = 4 * g + i + 16 * h;
                float2 c[R];
#pragma unroll
                for (int v = 0; v < R; ++v) {
                    c[v] = make_float2(yr[v][i], yi[v][i]);
                    if (v > 0) c[v] = c_mul(c[v], twv[v][kp]);
                }
                dftR<R>(c);
#pragma unroll
                for (int k2 = 0; k2 < R; ++k2) {
                    const float2 tw = k2 == 0 ? wk[h][i] : c_mul(wk[h][i], w32[k2]);
                    dst[m2 + (kp + 32 * k2) * M2] = c_mul(c[k2], tw);
                }
            }
        }
    }
}

// Shifts: frame f's integer shifts at shifts[2 f + station]; one launch
// per station, the kernel instance fixed by its code width (1..16 bits).
hipError_t launch_large_pass1_fused(const SectorArgs &a, const int64_t *shifts, float2 *out0,
                                    float2 *out1, int M1, int M2, const float2 *tw_m,
                                    hipStream_t s) {
    if (a.mf_tab == nullptr || M1 != 128 || (M2 != 64 && M2 != 128 && M2 != 256) || M1 * M2 != a.M)
        return hipErrorInvalidValue;
    if (a.nf <= 0) return hipSuccess;
    // 16-column blocks per frame: M2 / 16; one per wave, four waves per part
    const int parts = M2 / 64;
    for (int st = 0; st < 2; ++st) {
        if (a.st[st].bits < 1 || a.st[st].bits > 16) return hipErrorInvalidValue;
        const size_t lds = (size_t)(128 * (4 * a.st[st].bits + 1) + 1) * 4;   // <= 33,284 bytes
        auto go = [&](auto kern) -> hipError_t {
            hipLaunchKernelGGL(kern, dim3((unsigned)(a.nf * parts)), dim3(256), lds, s, a, st, shifts,
                               st == 0 ? out0 : out1, M2, parts, tw_m);
            return hipGetLastError();
        };
        hipError_t e;
        switch (a.st[st].bits) {
        case 1: e = go(k_large_pass1_fused<1>); break;
        case 2: e = go(k_large_pass1_fused<2>); break;
        case 4: e = go(k_large_pass1_fused<4>); break;
        case 8: e = go(k_large_pass1_fused<8>); break;
        default: e = go(k_large_pass1_fused<0>); break;
        }
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <typename IN>
static void launch_pass1_t(const IN *in, float2 *out, int64_t nbatch, int M1, int M2,
                           const uint4 *b, const float2 *tw_m, int64_t grid, hipStream_t s) {
    const int M = M1 * M2;
    if (M1 == 128)
        hipLaunchKernelGGL((k_fft_pass1_mfma<4, IN>), dim3((unsigned)grid), dim3(256), 0, s, in, out,
                           nbatch, M2, b, tw_m, M);
    else
        hipLaunchKernelGGL((k_fft_pass1_mfma<8, IN>), dim3((unsigned)grid), dim3(512), 0, s, in, out,
                           nbatch, M2, b, tw_m, M);
}

hipError_t launch_fft_pass1_mfma(const void *in, bool in_fp16, float2 *out, int64_t nbatch, int M1,
                                 int M2, const void *b1, const float2 *tw_m, hipStream_t s) {
    if (b1 == nullptr || M2 % 16 != 0 || (M1 != 128 && M1 != 256)) return hipErrorInvalidValue;
    if (nbatch <= 0) return hipSuccess;
    const int64_t grid = nbatch * (M2 / 16);
    const uint4 *b = reinterpret_cast<const uint4 *>(b1);
    if (in_fp16)
        launch_pass1_t(reinterpret_cast<const uint32_t *>(in), out, nbatch, M1, M2, b, tw_m, grid, s);
    else
        launch_pass1_t(reinterpret_cast<const float2 *>(in), out, nbatch, M1, M2, b, tw_m, grid, s);
    return hipGetLastError();
}

}  // namespace fxc
