// Chosen-form fused kernel for N = 1024 (M = 512 complex points per
// station-frame): four-step FFT whose first stage runs on the matrix cores.
//
//   z[m] = x[2m] + i x[2m+1],  m = 16*m1 + m2,  k = k1 + 32*k2
//   step 1 (MFMA, FP16 in / FP32 acc):  Y[k1][m2] = sum_m1 z[16 m1 + m2] W32^{m1 k1}
//   step 2 (FP32):                      Y'[k1][m2] = Y[k1][m2] W512^{m2 k1}
//   step 3 (FP32, radix-16):            Z[k1 + 32 k2] = sum_m2 Y'[k1][m2] W16^{m2 k2}
//
// Step 1 is a real GEMM D[m2][col] = sum_kk A[m2][kk] B[kk][col] with
// A = [Re z | Im z] (K = 64) and B = [[cos, -sin], [sin, cos]] (64 x 64
// outputs: Re / Im for k1 = 0..31).  The sample levels are exact in FP16;
// the DFT coefficients are split B = B_hi + 2^-11 B_lo (both FP16, B_lo
// scaled so it stays a normal number) with separate FP32 accumulators.
//
// Workgroup = 2 waves; wave s transforms station s for two frames per
// iteration (each MFMA's 16 rows are one frame's m2), so step 3 gives every
// lane one row.  The untangle to the real spectrum and the ACF/XCF products
// are the plain kernel's.
#include "fxc_dev.h"

namespace fxc {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int MF_M = 512;
constexpr int MF_YS = 36;                      // floats per transposed row (16 complex + pad)
constexpr int MF_BUF = 32 * MF_YS;             // Y rows; the padded Z spectrum aliases it
static_assert(MF_BUF >= 2 * (MF_M + MF_M / 16 + 2), "Z must fit in the Y buffer");
constexpr int MF_WORDS = 256 + 4;              // staged words of one station-frame (bits <= 8)

__device__ __forceinline__ half8 load_half8(const uint4 *p) {
    uint4 v = *p;
    half8 h;
    __builtin_memcpy(&h, &v, sizeof h);
    return h;
}

// One frame's A operand for K-step e (0: Re z, 1: Im z) of lane l: row
// m2 = l%16, K slots m1 = 8*(l/16) + j.  Element = sample 32*m1 + 2*m2 + e of
// the integer-shifted frame, as its FP16 level bits.
__device__ __forceinline__ half8 build_a(const uint32_t *words, const StationDev &st,
                                         const uint16_t *lev16, int l, int e, int sh, int N) {
    const int m2 = l & 15, g = l >> 4;
    const int base = 2 * m2 + e - sh;            // sample index minus 32*m1
    const int bb = base * st.bits;
    const int wq = bb >> 5;                      // floor(base*bits/32)
    const uint32_t off = (uint32_t)bb & 31u;
    const uint16_t sgn = (st.lsb && (base & 1)) ? 0x8000u : 0u;
    uint16_t v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int m1 = 8 * g + j;
        const int src = 32 * m1 + base;
        uint16_t h = 0;
        if ((unsigned)src < (unsigned)N) {
            const uint32_t code = (words[m1 * st.bits + wq] >> off) & st.code_mask;
            h = (uint16_t)(lev16[code] ^ sgn);
        }
        v[j] = h;
    }
    half8 r;
    __builtin_memcpy(&r, v, sizeof r);
    return r;
}

template <int PREC>
__global__ __launch_bounds__(128, 2) void k_sector_mfma(SectorArgs a) {
    __shared__ float buf[2][2][MF_BUF];          // [station][frame]: Y rows, then Z
    __shared__ uint32_t words[2][2][MF_WORDS];   // [station][frame]
    __shared__ FrameParam fp[kFrameBatch];
    __shared__ uint16_t lev16[2][256];
    __shared__ uint4 bsh[64 * 16];               // B operands [lane][hi 0..7 | lo 8..15]

    constexpr int M = MF_M, NT = 128, NB = (M + NT) / NT;
    constexpr int LOGNT = 7, LOGM = 9;
    const int tid = threadIdx.x;
    const int s = __builtin_amdgcn_readfirstlane(tid >> 6);   // station of this wave
    const int l = tid & 63;
    const int N = 2 * M;
    const StationDev &st = a.st[s];
    const uint8_t *raw = s ? a.raw2 : a.raw1;
    const int64_t f_begin = (int64_t)blockIdx.x * a.frames_per_wg;
    int64_t f_end = f_begin + a.frames_per_wg;
    if (f_end > a.nf) f_end = a.nf;

    for (int i = l; i < 256; i += 64) lev16[s][i] = (i <= (int)st.code_mask) ? a.lev16[s][i] : 0;

    for (int i = tid; i < 64 * 16; i += 128) bsh[i] = a.mfma_b[i];
    const uint4 *bl_ = bsh + l * 16;           // this lane's B operands
    const int k1 = l & 31, qrow = l >> 5;   // this lane's step-3 row

    double acc[4][NB];
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int i = 0; i < NB; ++i) acc[p][i] = 0.0;

    const bool do_acf = (a.flags & FLAG_ACF) != 0;
    const bool do_xcf = (a.flags & FLAG_XCF) != 0;
    const bool ov_only = (a.flags & FLAG_ACF_OVERLAP) != 0;
    const int gs = (a.grid == 0) ? a.a1s : a.a2s;
    const int ge = gs + a.ov;
    const int sb_other = (a.grid == 0) ? (a.a2s - a.a1s) : (a.a1s - a.a2s);
    const float lo_scale = 1.0f / 2048.0f;

    for (int64_t fb = f_begin; fb < f_end; fb += kFrameBatch) {
        const int nb = (int)((f_end - fb) < kFrameBatch ? (f_end - fb) : kFrameBatch);
        __syncthreads();
        if (tid < nb) make_frame_param<PREC>(a.first_frame + fb + tid, a, LOGM, fp[tid]);
        __syncthreads();
        for (int jf = 0; jf < nb; jf += 2) {
            const int nq = (nb - jf) >= 2 ? 2 : 1;
            // ---- stage this station's words for the (up to) two frames
#pragma unroll
            for (int q = 0; q < 2; ++q)
                if (q < nq) stage_words<64>(raw + (fb + jf + q) * st.bpf, st, words[s][q], l);
            __syncthreads();
            // ---- step 1 on the matrix cores, transposed into rows [k1][m2]
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int sh = (q < nq) ? (s ? fp[jf + q].s2 : fp[jf + q].s1) : N;
                const half8 a0 = build_a(words[s][q], st, lev16[s], l, 0, sh, N);
                const half8 a1 = build_a(words[s][q], st, lev16[s], l, 1, sh, N);
                float *yq = buf[s][q];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    f4 dh = {0.f, 0.f, 0.f, 0.f}, dl = {0.f, 0.f, 0.f, 0.f};
                    dh = __builtin_amdgcn_mfma_f32_16x16x32_f16(a0, load_half8(bl_ + t), dh, 0, 0, 0);
                    dh = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1, load_half8(bl_ + 4 + t), dh, 0, 0, 0);
                    dl = __builtin_amdgcn_mfma_f32_16x16x32_f16(a0, load_half8(bl_ + 8 + t), dl, 0, 0, 0);
                    dl = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1, load_half8(bl_ + 12 + t), dl, 0, 0, 0);
                    const int col = (l & 15) + 16 * (t & 1);   // k1
                    const int part = t >> 1;                   // 0: Re, 1: Im
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int m2 = 4 * (l >> 4) + i;
                        yq[col * MF_YS + 2 * m2 + part] = dh[i] + dl[i] * lo_scale;
                    }
                }
            }
            __syncthreads();
            // ---- steps 2 and 3: one row (k1, frame qrow) per lane
            {
                const float4 *row = reinterpret_cast<const float4 *>(buf[s][qrow] + k1 * MF_YS);
                float2 y[16];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const float4 v = row[j];
                    y[2 * j] = make_float2(v.x, v.y);
                    y[2 * j + 1] = make_float2(v.z, v.w);
                }
#pragma unroll
                for (int m2 = 1; m2 < 16; ++m2) y[m2] = c_mul(y[m2], a.tw_fft[(m2 * k1) & (M - 1)]);
                dft16(y);
                __syncthreads();                       // every row read before Z overwrites it
                float2 *zq = reinterpret_cast<float2 *>(buf[s][qrow]);
#pragma unroll
                for (int k2 = 0; k2 < 16; ++k2) zq[pad(k1 + 32 * k2)] = y[k2];
            }
            __syncthreads();
            // ---- products on the output grid (main.rs:2491-2568), per frame
            for (int q = 0; q < nq; ++q) {
                const FrameParam &P0 = fp[jf + q];
                const float2 *zg = reinterpret_cast<const float2 *>(buf[a.grid == 0 ? 0 : 1][q]);
                const float2 *zo = reinterpret_cast<const float2 *>(buf[a.grid == 0 ? 1 : 0][q]);
                const float2 frmix = P0.frmix;
                bool have = false;
                cd P = cd{0.0, 0.0};
#pragma unroll
                for (int i = 0; i < NB; ++i) {
                    const int g = tid + NT * i;
                    if (g <= M) {
                        const bool in_ov = (g >= gs) && (g < ge);
                        const float2 Xg = untangle(zg, g, M, a.tw_post);
                        const float pg = norm_sqr(Xg);
                        float2 Xo = make_float2(0.0f, 0.0f);
                        if (in_ov) Xo = untangle(zo, g + sb_other, M, a.tw_post);
                        if (do_acf) {
                            const float pgrid = (!ov_only || in_ov) ? pg : 0.0f;
                            const float poth = in_ov ? norm_sqr(Xo) : 0.0f;
                            acc[0][i] += (double)(a.grid == 0 ? pgrid : poth);
                            acc[1][i] += (double)(a.grid == 0 ? poth : pgrid);
                        }
                        if (in_ov && do_xcf) {
                            const int k = g - gs;
                            if (!have) {
                                P = P0.start;
                                // first overlap bin of this thread: k < NT
#pragma unroll 1
                                for (int b = 0; b < LOGNT; ++b)
                                    if ((k >> b) & 1) P = d_mul(P, P0.pw[b]);
                                have = true;
                            } else {
                                P = d_mul(P, P0.pw[LOGNT]);
                            }
                            const float2 ph = make_float2((float)P.x, (float)P.y);
                            const float2 X1 = (a.grid == 0) ? Xg : Xo;
                            const float2 X2 = (a.grid == 0) ? Xo : Xg;
                            float2 v2 = c_mul(X1, c_conj(X2));
                            if (PREC == 0) v2 = c_mul(v2, frmix);
                            v2 = c_mul(v2, ph);
                            acc[2][i] += (double)v2.x;
                            acc[3][i] += (double)v2.y;
                        }
                    }
                }
            }
            __syncthreads();
        }
    }

    const int half = M + 1;
    double *out = a.part + (size_t)blockIdx.x * kNumProducts * half;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int g = tid + NT * i;
        if (g <= M) {
            out[P_PH * half + g] = 0.0;
            out[P_11 * half + g] = acc[0][i];
            out[P_22 * half + g] = acc[1][i];
            out[P_12R * half + g] = acc[2][i];
            out[P_12I * half + g] = acc[3][i];
        }
    }
}

hipError_t launch_sector_mfma(const SectorArgs &a, int precision, int n_wg, hipStream_t s) {
    if (a.N != 2 * MF_M || a.mfma_b == nullptr) return hipErrorInvalidValue;
    if (precision == 0)
        hipLaunchKernelGGL(k_sector_mfma<0>, dim3(n_wg), dim3(128), 0, s, a);
    else
        hipLaunchKernelGGL(k_sector_mfma<1>, dim3(n_wg), dim3(128), 0, s, a);
    return hipGetLastError();
}

// Forward R2C of frames through the MFMA four-step FFT (test entry): two
// frames per wave, one wave per workgroup.
__global__ __launch_bounds__(64) void k_rfft_mfma(SectorArgs a, const float *in, int64_t nframes,
                                                  float2 *out) {
    __shared__ float buf[2][MF_BUF];
    constexpr int M = MF_M;
    const int l = threadIdx.x;
    half8 bh[2][4], bl[2][4];
#pragma unroll
    for (int e = 0; e < 2; ++e)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            bh[e][t] = load_half8(a.mfma_b + l * 16 + e * 4 + t);
            bl[e][t] = load_half8(a.mfma_b + l * 16 + 8 + e * 4 + t);
        }
    const int k1 = l & 31, qrow = l >> 5;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int64_t f = 2 * (int64_t)blockIdx.x + q;
        const int m2 = l & 15, g = l >> 4;
        _Float16 va[8], vb[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int m1 = 8 * g + j;
            const float x0 = (f < nframes) ? in[f * 2 * M + 32 * m1 + 2 * m2] : 0.f;
            const float x1 = (f < nframes) ? in[f * 2 * M + 32 * m1 + 2 * m2 + 1] : 0.f;
            va[j] = (_Float16)x0;
            vb[j] = (_Float16)x1;
        }
        half8 a0, a1;
        __builtin_memcpy(&a0, va, sizeof a0);
        __builtin_memcpy(&a1, vb, sizeof a1);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            f4 dh = {0.f, 0.f, 0.f, 0.f}, dl = {0.f, 0.f, 0.f, 0.f};
            dh = __builtin_amdgcn_mfma_f32_16x16x32_f16(a0, bh[0][t], dh, 0, 0, 0);
            dh = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1, bh[1][t], dh, 0, 0, 0);
            dl = __builtin_amdgcn_mfma_f32_16x16x32_f16(a0, bl[0][t], dl, 0, 0, 0);
            dl = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1, bl[1][t], dl, 0, 0, 0);
            const int col = (l & 15) + 16 * (t & 1);
            const int part = t >> 1;
#pragma unroll
            for (int i = 0; i < 4; ++i)
                buf[q][col * MF_YS + 2 * (4 * (l >> 4) + i) + part] = dh[i] + dl[i] * (1.0f / 2048.0f);
        }
    }
    __syncthreads();
    float2 y[16];
    const float4 *row = reinterpret_cast<const float4 *>(buf[qrow] + k1 * MF_YS);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const float4 v = row[j];
        y[2 * j] = make_float2(v.x, v.y);
        y[2 * j + 1] = make_float2(v.z, v.w);
    }
#pragma unroll
    for (int m2 = 1; m2 < 16; ++m2) y[m2] = c_mul(y[m2], a.tw_fft[(m2 * k1) & (M - 1)]);
    dft16(y);
    __syncthreads();
    float2 *zq = reinterpret_cast<float2 *>(buf[qrow]);
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) zq[pad(k1 + 32 * k2)] = y[k2];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int64_t f = 2 * (int64_t)blockIdx.x + q;
        if (f < nframes)
            for (int k = l; k <= M; k += 64)
                out[f * (M + 1) + k] = untangle(reinterpret_cast<const float2 *>(buf[q]), k, M, a.tw_post);
    }
}

hipError_t launch_rfft_mfma(const SectorArgs &a, const float *in, int64_t nframes, float2 *out,
                            hipStream_t s) {
    if (a.N != 2 * MF_M || a.mfma_b == nullptr) return hipErrorInvalidValue;
    if (nframes <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_rfft_mfma, dim3((unsigned)((nframes + 1) / 2)), dim3(64), 0, s, a, in,
                       nframes, out);
    return hipGetLastError();
}

}  // namespace fxc
