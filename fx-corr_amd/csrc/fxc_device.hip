// Device code of the FX hot path for gfx950 (MI355X).
//
// One workgroup owns a contiguous run of frames of one sector.  For every
// frame pair it: stages the packed words of both stations in LDS (shuffle
// applied once per word), decodes + integer-shifts straight into the first
// FFT pass's registers, runs an LDS Stockham complex FFT of N/2 points per
// station, untangles the real spectrum (realfft's method), and accumulates
// ACF/XCF products into FP64 registers owned per output bin.  Partials are
// written once per workgroup and reduced in a fixed order (deterministic).
#include "fxc_dev.h"

namespace fxc {

// The workgroup holds 2T threads: threads [0,T) decode and transform
// station 1, [T,2T) station 2 (M = E*T complex points each); then every
// thread owns output-grid bins g = tid + 2T*i for the products.
// MF (chosen form, levels exact in FP16): the FFT's first pass (radix 32)
// runs as FP16 hi/lo MFMA GEMMs, the remaining Stockham passes in FP32.
// The MF instance at E = 8, T = 256 is held to 128 VGPRs (four waves per
// SIMD, two 512-thread workgroups per CU); unconstrained it takes 130 and
// one workgroup per CU.
template <int E, int T, int PREC, bool MF = false>
__global__ __launch_bounds__(2 * T)
__attribute__((amdgpu_waves_per_eu(MF && E == 8 && T == 256 ? 4 : 1))) void k_sector(SectorArgs a) {
    constexpr int M = E * T;
    constexpr int MP = M + M / 16 + 2;
    constexpr int NT = 2 * T;
    constexpr int NB = (M + NT) / NT;            // ceil((M+1)/NT)
    constexpr int LOGNT = __builtin_ctz(NT);
    constexpr int LOGM = __builtin_ctz(M);
    static_assert(LOGM <= kMaxPow && LOGNT < kMaxPow, "phase power table too small");
    __shared__ float2 zbuf[2][MP];
    __shared__ FrameParam fp[kFrameBatch];
    __shared__ float lev[2][256];
    // MF: the DFT-32 operand table in LDS (16 KB; at M = 8192 the kernel's
    // LDS is then 160,032 of 163,840 bytes)
    constexpr bool B1_LDS = MF;
    __shared__ uint4 b1s[B1_LDS ? 16 * 64 : 1];
    // plain-form chain prefetch at N = 8192, where it also relieves the
    // register spills; at N = 2048/4096 its registers cost a wave per SIMD
    // and the kernel got slower
    constexpr bool CPF = PREC == 0 && NB <= 5 && T >= 512;

    const int tid0 = threadIdx.x;
    const int s = half_index<T>(tid0);   // station handled by this half of the workgroup
    const int t0 = tid0 - s * T;
    const int tid = tid0, t = t0;
    const int N = 2 * M;
    const int64_t f_begin = (int64_t)blockIdx.x * a.frames_per_wg;
    int64_t f_end = f_begin + a.frames_per_wg;
    if (f_end > a.nf) f_end = a.nf;
    const StationDev &st = a.st[s];
    const uint8_t *raw = s ? a.raw2 : a.raw1;

    if (st.bits > 2 && st.bits <= 8)
        for (int i = t; i < 256; i += T) lev[s][i] = (i <= (int)st.code_mask) ? st.levels[i] : 0.0f;
    if constexpr (B1_LDS)
        for (int i = tid0; i < 16 * 64; i += NT) b1s[i] = reinterpret_cast<const uint4 *>(a.mf_tab)[i];
    const uint4 *b1 = B1_LDS ? b1s : reinterpret_cast<const uint4 *>(a.mf_tab);

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
    float2 *zs = zbuf[s];
    uint32_t *ws = reinterpret_cast<uint32_t *>(zbuf[s]);
    const float2 *zg = zbuf[a.grid == 0 ? 0 : 1];
    const float2 *zo = zbuf[a.grid == 0 ? 1 : 0];
    __syncthreads();

    for (int64_t fb = f_begin; fb < f_end; fb += kFrameBatch) {
        const int nb = (int)((f_end - fb) < kFrameBatch ? (f_end - fb) : kFrameBatch);
        if (tid < nb) make_frame_param<PREC>(a.first_frame + fb + tid, a, LOGM, fp[tid]);
        __syncthreads();
        for (int jf = 0; jf < nb; ++jf) {
            const int64_t f = fb + jf;
            const int tid = loop_local(tid0), t = loop_local(t0);   // see loop_local
            stage_words<T>(raw + f * st.bpf, st, ws, t);
            __syncthreads();
            const int sh = s ? fp[jf].s2 : fp[jf].s1;
            // plain form: the frame's chain values (the reference's own f32
            // XCF phasors) loaded before the decode and FFT so their latency
            // overlaps them
            float2 chv[CPF ? NB : 1];
            if constexpr (CPF) {
                if (a.chain != nullptr) {
#pragma unroll
                    for (int i = 0; i < NB; ++i) {
                        const int g = tid + NT * i;
                        chv[i] = (g <= M && g >= gs && g < ge) ? a.chain[f * a.chain_ld + (g - gs)]
                                                               : make_float2(0.f, 0.f);
                    }
                }
            }
            float2 v[E];
            if constexpr (MF) {
                // the frame goes to LDS before the first pass: each thread
                // decodes a run of E consecutive elements
                decode_run<E>(ws, st, lev[s], t * E, sh, N, v);
                __syncthreads();
#pragma unroll
                for (int r = 0; r < E; ++r) zs[pad(t * E + r)] = v[r];
                __syncthreads();
                stockham_first_mfma<E, T>(zs, b1, t);
                stockham_rest<E, T, 32>(zs, a.tw_fft, t);
            } else {
#pragma unroll
                for (int r = 0; r < E; ++r) v[r] = decode_elem(ws, st, lev[s], t + T * r, sh, N);
                __syncthreads();
                fft_forward<E, T>(zs, v, a.tw_fft, t);
            }

            // ---- products on the output grid (main.rs:2491-2568)
            const float2 frmix = fp[jf].frmix;
            bool have = false;
            cd P = cd{0.0, 0.0};
            float2 Pf = make_float2(0.f, 0.f), Ff = Pf;
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
                        // grid station: every bin (or overlap only); other station: overlap
                        const float pgrid = (!ov_only || in_ov) ? pg : 0.0f;
                        const float poth = in_ov ? norm_sqr(Xo) : 0.0f;
                        acc[0][i] += (double)(a.grid == 0 ? pgrid : poth);
                        acc[1][i] += (double)(a.grid == 0 ? poth : pgrid);
                    }
                    if (in_ov && do_xcf) {
                        const int k = g - gs;
                        float2 ph;
                        if (PREC == 0 && a.chain != nullptr) {
                            ph = CPF ? chv[CPF ? i : 0] : a.chain[f * a.chain_ld + k];   // the reference's own chain
                        } else if (PREC != 0) {
                            // chosen form: the first overlap bin's phasor from
                            // the FP64 phase in turns (reduced mod 1, then the
                            // f32 sincos), the thread's later bins (k + NT)
                            // by FP32 steps of exp(2 pi i NT dth)
                            if (!have) {
                                const float r = (float)wrap_turns(fma((double)k, fp[jf].dth, fp[jf].phw));
                                Pf = make_float2(__builtin_amdgcn_cosf(r), __builtin_amdgcn_sinf(r));
                                const float rs = (float)wrap_turns((double)NT * fp[jf].dth);
                                Ff = make_float2(__builtin_amdgcn_cosf(rs), __builtin_amdgcn_sinf(rs));
                                have = true;
                            } else {
                                Pf = c_mul(Pf, Ff);
                            }
                            ph = Pf;
                        } else {
                            if (!have) {
                                P = fp[jf].start;
                                // first overlap bin of this thread: k < NT
#pragma unroll 1
                                for (int b = 0; b < LOGNT; ++b)
                                    if ((k >> b) & 1) P = d_mul(P, fp[jf].pw[b]);
                                have = true;
                            } else {
                                P = d_mul(P, fp[jf].pw[LOGNT]);
                            }
                            ph = make_float2((float)P.x, (float)P.y);
                        }
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
            __syncthreads();
        }
    }

    // ---- workgroup partials [wg][ph|11|22|12re|12im][half]
    const int half = M + 1;
    double *out = a.part + (size_t)blockIdx.x * kNumProducts * half;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int g = tid + NT * i;
        if (g <= M) {
            const bool add = a.accumulate != 0;
            out[P_PH * half + g] = 0.0;
            out[P_11 * half + g] = add ? out[P_11 * half + g] + acc[0][i] : acc[0][i];
            out[P_22 * half + g] = add ? out[P_22 * half + g] + acc[1][i] : acc[1][i];
            out[P_12R * half + g] = add ? out[P_12R * half + g] + acc[2][i] : acc[2][i];
            out[P_12I * half + g] = add ? out[P_12I * half + g] + acc[3][i] : acc[3][i];
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

// The same reduce over partial slots held in the fused four-step's
// pair-major order (index p 2 M2 + j: row p for j < M2, row M1 - p, or
// M1 / 2 for p = 0, otherwise; k2 = j mod M2; bin M last), written to the
// bins g = row + M1 k2.
__global__ void k_reduce_final_pairs(const double *scratch, int n_chunks, int half, int M1,
                                     int M2, double *out) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= kNumProducts * half) return;
    double s = 0.0;
    for (int c = 0; c < n_chunks; ++c) s += scratch[(size_t)c * kNumProducts * half + col];
    const int prod = col / half, idx = col - prod * half;
    int g = idx;
    if (idx < half - 1) {
        const int p = idx / (2 * M2), j = idx - p * 2 * M2;
        const int row = j < M2 ? p : (p == 0 ? M1 / 2 : M1 - p);
        g = row + M1 * (j % M2);
    }
    out[(size_t)prod * half + g] = s;
}

hipError_t launch_reduce_pairs(const double *part, int n_wg, int half, int M1, int M2, double *out,
                               double *scratch, int n_chunks, hipStream_t stream) {
    const int cols = kNumProducts * half;
    if (n_chunks < 1) n_chunks = 1;
    if (n_chunks > n_wg) n_chunks = n_wg;
    const int chunk = (n_wg + n_chunks - 1) / n_chunks;
    n_chunks = (n_wg + chunk - 1) / chunk;
    dim3 g1((cols + 255) / 256, n_chunks);
    hipLaunchKernelGGL(k_reduce_chunks, g1, dim3(256), 0, stream, part, n_wg, cols, chunk, scratch);
    hipLaunchKernelGGL(k_reduce_final_pairs, dim3((cols + 255) / 256), dim3(256), 0, stream, scratch,
                       n_chunks, half, M1, M2, out);
    return hipGetLastError();
}

// ---------------------------------------------------------- dispatch
// Elements per thread E and threads per station T for a given N
// (M = N/2 = E*T; the workgroup holds 2T threads).
static void shape_for(int N, int &E, int &T) {
    const int M = N / 2;
    if (M >= 8192) {
        E = 16;
        T = M / 16;
    } else if (M >= 512) {
        E = 8;
        T = M / 8;
    } else if (M == 256) {
        E = 4;
        T = 64;
    } else if (M == 128) {   // one wave per frame pair: radix-4 passes (129k vs 101k
        E = 4;               // Msamples/s/station with two waves of radix-2 passes)
        T = 32;
    } else {
        E = 2;
        T = M / 2;
    }
}

bool sector_supported(int N) { return N >= 64 && N <= 16384 && (N & (N - 1)) == 0; }
int sector_threads_per_frame(int N) {
    int E, T;
    shape_for(N, E, T);
    return 2 * T;
}

template <int E, int T>
static hipError_t launch_sector_t(const SectorArgs &a, int precision, int n_wg, hipStream_t s) {
    if (precision == 0) {
        hipLaunchKernelGGL((k_sector<E, T, 0>), dim3(n_wg), dim3(2 * T), 0, s, a);
        return hipGetLastError();
    }
    // a.mf_tab: the plan's first FFT pass runs on FP16 MFMA (whole waves of
    // 32-point DFT rows: E >= 8, T a multiple of 64)
    if constexpr (E >= 8 && T % 64 == 0) {
        if (a.mf_tab != nullptr) {
            hipLaunchKernelGGL((k_sector<E, T, 1, true>), dim3(n_wg), dim3(2 * T), 0, s, a);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((k_sector<E, T, 1>), dim3(n_wg), dim3(2 * T), 0, s, a);
    return hipGetLastError();
}

// a.variant 1 selects radix-16 passes with half the threads per station
// (N = 2048..8192): a different FFT split, chosen by the plan autotuner.
hipError_t launch_sector(const SectorArgs &a, int precision, int n_wg, hipStream_t s) {
    const bool v16 = a.variant == 1;
    switch (a.N) {
    case 64: return launch_sector_t<2, 16>(a, precision, n_wg, s);
    case 128: return launch_sector_t<2, 32>(a, precision, n_wg, s);
    case 256: return launch_sector_t<4, 32>(a, precision, n_wg, s);
    case 512: return launch_sector_t<4, 64>(a, precision, n_wg, s);
    case 1024: return launch_sector_t<8, 64>(a, precision, n_wg, s);
    case 2048: return v16 ? launch_sector_t<16, 64>(a, precision, n_wg, s)
                          : launch_sector_t<8, 128>(a, precision, n_wg, s);
    case 4096: return v16 ? launch_sector_t<16, 128>(a, precision, n_wg, s)
                          : launch_sector_t<8, 256>(a, precision, n_wg, s);
    case 8192: return v16 ? launch_sector_t<16, 256>(a, precision, n_wg, s)
                          : launch_sector_t<8, 512>(a, precision, n_wg, s);
    case 16384: return launch_sector_t<16, 512>(a, precision, n_wg, s);
    default: return hipErrorInvalidValue;
    }
}

bool sector_has_variant(int N) { return N >= 2048 && N <= 8192; }

// ------------------------------------------ per-function entry kernels
// Decode (+ shift) of one frame per workgroup into out[f*N + n], through the
// fused kernel's staging and element decode.  H: each element (x[2m],
// x[2m+1]) as an FP16 pair in one dword (exact for levels that are exact in
// FP16), else as two floats.
template <int T, bool H>
__global__ __launch_bounds__(T) void k_decode_frames(SectorArgs a, int station,
                                                     const int64_t *shifts, int shift_stride,
                                                     void *out) {
    extern __shared__ uint32_t words[];   // st.nwords
    __shared__ float lev[256];
    const StationDev &st = a.st[station];
    const int t = threadIdx.x;
    const int64_t f = blockIdx.x;
    if (st.bits > 2 && st.bits <= 8)
        for (int i = t; i < 256; i += T) lev[i] = (i <= (int)st.code_mask) ? st.levels[i] : 0.0f;
    const uint8_t *raw = (station == 0 ? a.raw1 : a.raw2) + f * st.bpf;
    stage_words<T>(raw, st, words, t);
    __syncthreads();
    const int s = shifts ? clamp_shift(shifts[f * shift_stride], a.N) : 0;
    for (int m = t; m < a.M; m += T) {
        const float2 v = decode_elem(words, st, lev, m, s, a.N);
        if constexpr (H)
            reinterpret_cast<uint32_t *>(out)[f * a.M + m] =
                __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(v.x, v.y));
        else
            reinterpret_cast<float2 *>(out)[f * a.M + m] = v;
    }
}

template <int T, bool H>
static hipError_t launch_decode_frames_t(const SectorArgs &a, int station, const int64_t *shifts,
                                         int shift_stride, void *out, hipStream_t s) {
    const size_t lds = (size_t)a.st[station].nwords * 4;
    if (lds > 96 * 1024) return hipErrorInvalidValue;
    if (lds > 48 * 1024) {   // above the default dynamic-LDS limit (8-bit samples, N = 65536)
        hipError_t e = hipFuncSetAttribute((const void *)k_decode_frames<T, H>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_decode_frames<T, H>), dim3((unsigned)a.nf), dim3(T), lds, s, a, station,
                       shifts, shift_stride, out);
    return hipGetLastError();
}

// N >= 16384: 1024 threads per frame.  The HBM four-step path decodes chunks
// of a few hundred frames, and 256-thread workgroups left about two waves per
// SIMD.
template <bool H>
static hipError_t launch_decode_frames_h(const SectorArgs &a, int station, const int64_t *shifts,
                                         int shift_stride, void *out, hipStream_t s) {
    return a.M >= 8192 ? launch_decode_frames_t<1024, H>(a, station, shifts, shift_stride, out, s)
                       : launch_decode_frames_t<256, H>(a, station, shifts, shift_stride, out, s);
}

hipError_t launch_decode_frames(const SectorArgs &a, int station, const int64_t *shifts,
                                int shift_stride, void *out, bool fp16, hipStream_t s) {
    if (a.nf <= 0) return hipSuccess;
    return fp16 ? launch_decode_frames_h<true>(a, station, shifts, shift_stride, out, s)
                : launch_decode_frames_h<false>(a, station, shifts, shift_stride, out, s);
}

// Forward R2C with the fused kernel's FFT code: each half of the workgroup
// transforms one frame (MF: first pass on FP16 MFMA, as k_sector).
template <int E, int T, bool MF = false>
__global__ __launch_bounds__(2 * T) void k_rfft_frames(SectorArgs a, const float *in,
                                                       int64_t nframes, float2 *out) {
    constexpr int M = E * T;
    constexpr int MP = M + M / 16 + 2;
    __shared__ float2 zbuf[2][MP];
    const int s = half_index<T>(threadIdx.x);
    const int t = threadIdx.x - s * T;
    const int64_t f = 2 * (int64_t)blockIdx.x + s;
    const bool live = f < nframes;
    float2 v[E];
#pragma unroll
    for (int r = 0; r < E; ++r)
        v[r] = live ? reinterpret_cast<const float2 *>(in + f * 2 * M)[t + T * r]
                    : make_float2(0.f, 0.f);
    if constexpr (MF) {
#pragma unroll
        for (int r = 0; r < E; ++r) zbuf[s][pad(t + T * r)] = v[r];
        __syncthreads();
        stockham_first_mfma<E, T>(zbuf[s], reinterpret_cast<const uint4 *>(a.mf_tab), t);
        stockham_rest<E, T, 32>(zbuf[s], a.tw_fft, t);
    } else {
        fft_forward<E, T>(zbuf[s], v, a.tw_fft, t);
    }
    if (live)
        for (int k = t; k <= M; k += T) out[f * (M + 1) + k] = untangle(zbuf[s], k, M, a.tw_post);
}

// Inverse C2R (scaled 1/N): build Z' = E' + i O' from the half spectrum,
// IFFT via conj(FFT(conj(.))), interleave, scale.  DC/Nyquist imaginary
// parts are ignored (the reference zeroes them first).
template <int E, int T>
__device__ __forceinline__ void irfft_prepare(const float2 *X, int M, const float2 *tw_post,
                                              float2 *v, int t) {
#pragma unroll
    for (int r = 0; r < E; ++r) v[r] = irfft_pre_elem(X, t + T * r, M, tw_post);
}

template <int E, int T>
__global__ __launch_bounds__(2 * T) void k_irfft_frames(SectorArgs a, const float2 *in,
                                                        int64_t nframes, float *out,
                                                        int negate_odd) {
    constexpr int M = E * T;
    constexpr int MP = M + M / 16 + 2;
    __shared__ float2 zbuf[2][MP];
    const int s = half_index<T>(threadIdx.x);
    const int t = threadIdx.x - s * T;
    const int64_t f = 2 * (int64_t)blockIdx.x + s;
    const bool live = f < nframes;
    float2 v[E];
    if (live) irfft_prepare<E, T>(in + f * (M + 1), M, a.tw_post, v, t);
    else
        for (int r = 0; r < E; ++r) v[r] = make_float2(0.f, 0.f);
    fft_forward<E, T>(zbuf[s], v, a.tw_fft, t);
    const float scale = 1.0f / (float)(2 * M);
    // odd samples are -Im of the conjugated transform, negated once more for
    // an LSB output stream (main.rs:2290-2292)
    const float odd_sign = negate_odd ? 1.0f : -1.0f;
    if (live)
        for (int m = t; m < M; m += T) {
            float2 z = zbuf[s][pad(m)];
            out[f * 2 * M + 2 * m] = z.x * scale;
            out[f * 2 * M + 2 * m + 1] = odd_sign * (z.y * scale);
        }
}

// mode 0: forward R2C; 1: inverse C2R; 2: inverse with odd samples negated.
template <int E, int T>
static hipError_t launch_fft_t(const SectorArgs &a, const void *in, int64_t nframes, void *out,
                               int mode, hipStream_t s) {
    dim3 grid((unsigned)((nframes + 1) / 2));
    if (mode != 0)
        hipLaunchKernelGGL((k_irfft_frames<E, T>), grid, dim3(2 * T), 0, s, a, (const float2 *)in,
                           nframes, (float *)out, mode == 2 ? 1 : 0);
    else {
        if constexpr (E >= 8 && T % 64 == 0) {
            if (a.mf_tab != nullptr) {
                hipLaunchKernelGGL((k_rfft_frames<E, T, true>), grid, dim3(2 * T), 0, s, a,
                                   (const float *)in, nframes, (float2 *)out);
                return hipGetLastError();
            }
        }
        hipLaunchKernelGGL((k_rfft_frames<E, T>), grid, dim3(2 * T), 0, s, a, (const float *)in,
                           nframes, (float2 *)out);
    }
    return hipGetLastError();
}

static hipError_t launch_fft(const SectorArgs &a, const void *in, int64_t nframes, void *out,
                             int mode, hipStream_t s) {
    if (nframes <= 0) return hipSuccess;
    switch (a.N) {
    case 64: return launch_fft_t<2, 16>(a, in, nframes, out, mode, s);
    case 128: return launch_fft_t<2, 32>(a, in, nframes, out, mode, s);
    case 256: return launch_fft_t<4, 32>(a, in, nframes, out, mode, s);
    case 512: return launch_fft_t<4, 64>(a, in, nframes, out, mode, s);
    case 1024: return launch_fft_t<8, 64>(a, in, nframes, out, mode, s);
    case 2048: return launch_fft_t<8, 128>(a, in, nframes, out, mode, s);
    case 4096: return launch_fft_t<8, 256>(a, in, nframes, out, mode, s);
    case 8192: return launch_fft_t<8, 512>(a, in, nframes, out, mode, s);
    case 16384: return launch_fft_t<16, 512>(a, in, nframes, out, mode, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_rfft_frames(const SectorArgs &a, const float *in, int64_t nframes, float2 *out,
                              hipStream_t s) {
    return launch_fft(a, in, nframes, out, 0, s);
}
hipError_t launch_irfft_frames(const SectorArgs &a, const float2 *in, int64_t nframes, float *out,
                               bool negate_odd, hipStream_t s) {
    return launch_fft(a, in, nframes, out, negate_odd ? 2 : 1, s);
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

// The same code for non-decreasing levels, from lo, the last level <= x
// (-1: none).  |fl(x - l_k)| is non-increasing in k up to lo and
// non-decreasing after it (rounding is monotone), so the scan's first
// minimum is lo + 1 when that error is strictly smaller, else the first of
// the run of errors equal to level lo's.  As in the scan, an error that does
// not compare below FLT_MAX (x infinite) leaves code 0.
__device__ __forceinline__ uint32_t quant_code_from_lo(float x, const float *lv, int nl, int lo) {
    const int hi = lo + 1;
    int best;
    float eb;
    if (lo < 0) {
        best = 0;
        eb = fabsf(x - lv[0]);
    } else {
        best = lo;
        eb = fabsf(x - lv[lo]);
        const float eh = hi < nl ? fabsf(x - lv[hi]) : eb;
        if (eh < eb) {
            best = hi;
            eb = eh;
        } else {
            while (best > 0 && fabsf(x - lv[best - 1]) == eb) --best;
        }
    }
    return eb < 3.40282347e+38f ? (uint32_t)best : 0u;
}

// Non-decreasing levels: lo by bisection.  x NaN: code 0, as the scan.
__device__ __forceinline__ uint32_t quant_code_sorted(float x, const float *lv, int nl) {
    if (!(x == x)) return 0u;
    int lo = -1, hi = nl;   // lv[lo] <= x < lv[hi] (lo = -1, hi = nl: none)
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (lv[mid] <= x) lo = mid;
        else hi = mid;
    }
    return quant_code_from_lo(x, lv, nl, lo);
}

// Output quantiser levels and the search that finds lo:
// mode 0 the scan (any level order), 1 bisection (non-decreasing levels),
// 2 (non-decreasing levels within a quarter step of an even spacing) a
// guess floor((x - l0) inv) moved to the last level <= x by comparisons,
// the same lo as the bisection's in about one step.
struct QLev {
    const float *lv;
    int nl;
    int mode;
    float l0, inv;
};

// Mode 2 fast path: the four levels k0 .. k0 + 3 around the guess (k0 =
// g - 1, clamped into the levels).  When they bracket x (lv[k0] <= x or
// k0 = 0; lv[k0 + 3] > x or k0 + 3 the last level) they hold lo and lo + 1,
// and the errors are non-increasing up to lo and non-decreasing after it, so
// the scan's first minimum is the window's first minimum, unless that is
// level k0 and level k0 - 1 ties with it (then the general path walks the
// tie run).  Four independent LDS reads, no data-dependent loop.
__device__ __forceinline__ uint32_t quant_code_q(float x, const QLev &q) {
    if (q.mode == 0) return quant_code(x, q.lv, q.nl);
    if (q.mode == 1) return quant_code_sorted(x, q.lv, q.nl);
    if (!(x == x)) return 0u;
    const float gf = fminf(fmaxf(floorf((x - q.l0) * q.inv), -1.0f), (float)(q.nl - 1));
    int g = (int)gf;
    if (q.nl >= 4) {
        const int k0 = min(max(g - 1, 0), q.nl - 4);
        const float v0 = q.lv[k0], v1 = q.lv[k0 + 1], v2 = q.lv[k0 + 2], v3 = q.lv[k0 + 3];
        if ((k0 == 0 || v0 <= x) && (k0 + 3 == q.nl - 1 || v3 > x)) {
            float eb = fabsf(x - v0);
            int best = k0;
            const float e1 = fabsf(x - v1), e2 = fabsf(x - v2), e3 = fabsf(x - v3);
            if (e1 < eb) { eb = e1; best = k0 + 1; }
            if (e2 < eb) { eb = e2; best = k0 + 2; }
            if (e3 < eb) { eb = e3; best = k0 + 3; }
            if (!(best == k0 && k0 > 0 && fabsf(x - q.lv[k0 - 1]) == eb))
                return eb < 3.40282347e+38f ? (uint32_t)best : 0u;
        }
    }
    while (g + 1 < q.nl && q.lv[g + 1] <= x) ++g;
    while (g >= 0 && q.lv[g] > x) --g;
    return quant_code_from_lo(x, q.lv, q.nl, g);
}

__device__ __forceinline__ QLev out_qlev(const SectorArgs &a, const float *lv_lds) {
    QLev q;
    q.lv = lv_lds;
    q.nl = a.out_nlevels < 256 ? a.out_nlevels : 256;
    q.mode = a.out_qmode;
    q.l0 = a.out_l0;
    q.inv = a.out_inv_step;
    return q;
}

__device__ __forceinline__ uint32_t shuffle_out_word(uint32_t w, const int32_t *sh) {
    uint32_t o = 0;
#pragma unroll
    for (int p = 0; p < 32; ++p) o |= ((w >> p) & 1u) << sh[p];
    return o;
}

// One thread per output word of one frame; x(s) gives sample s.
template <typename X>
__device__ __forceinline__ uint32_t quant_word(X x, int N, int w, int bits, const QLev &ql,
                                               uint32_t mask) {
    uint64_t acc = 0;
    const int64_t b0 = (int64_t)w * 32, b1 = b0 + 32;
    int64_t s = b0 / bits;
    for (; s * bits < b1 && s < N; ++s) {
        const float xs = x((int)s);
        const uint32_t q = quant_code_q(xs, ql);
        uint64_t code = (uint64_t)(q & mask);
        int64_t pos = s * bits - b0;  // may be negative for the straddling first sample
        if (pos >= 0) acc |= code << pos;
        else acc |= code >> (-pos);
    }
    return (uint32_t)(acc & 0xFFFFFFFFull);
}

// The same word when the code width divides 32 (1, 2, 4, 8, 16, 32 bits):
// samples w spw .. w spw + spw - 1, no code straddles a word, 32-bit
// arithmetic only.
template <typename X>
__device__ __forceinline__ uint32_t quant_word_aligned(X x, int w, int bits, const QLev &ql,
                                                       uint32_t mask) {
    const int spw = 32 / bits;
    uint32_t acc = 0;
    for (int k = 0; k < spw; ++k) {
        const float xs = x(w * spw + k);
        const uint32_t q = quant_code_q(xs, ql);
        acc |= (q & mask) << (k * bits);
    }
    return acc;
}

// The aligned word at a compile-time width: the sample loop unrolled, so
// the two samples of one complex value share its LDS read.
template <int BITS, typename X>
__device__ __forceinline__ uint32_t quant_word_bits(X x, int w, const QLev &ql) {
    constexpr int spw = 32 / BITS;
    constexpr uint32_t mask = BITS >= 32 ? 0xFFFFFFFFu : ((1u << BITS) - 1u);
    uint32_t acc = 0;
#pragma unroll
    for (int k = 0; k < spw; ++k) acc |= (quant_code_q(x(w * spw + k), ql) & mask) << (k * BITS);
    return acc;
}

// Word w of a frame of N samples at any width: the common widths unrolled.
template <typename X>
__device__ __forceinline__ uint32_t quant_word_any(X x, int N, int w, int bits, const QLev &ql,
                                                   uint32_t mask) {
    switch (bits) {
    case 1: return quant_word_bits<1>(x, w, ql);
    case 2: return quant_word_bits<2>(x, w, ql);
    case 4: return quant_word_bits<4>(x, w, ql);
    case 8: return quant_word_bits<8>(x, w, ql);
    case 16: return quant_word_bits<16>(x, w, ql);
    default:
        return (32 % bits) == 0 ? quant_word_aligned(x, w, bits, ql, mask)
                                : quant_word(x, N, w, bits, ql, mask);
    }
}

// The output shuffle as four byte tables in LDS: bit 8 b + j of a word goes
// to bit out_shifts[8 b + j] (shuffle_out_word's map).
__device__ __forceinline__ void build_out_shuffle(uint32_t (*tab)[256], const int32_t *sh, int tid,
                                                  int nt) {
    for (int i = tid; i < 4 * 256; i += nt) {
        const int b = i >> 8, v = i & 255;
        uint32_t o = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) o |= ((uint32_t)(v >> j) & 1u) << sh[8 * b + j];
        tab[b][v] = o;
    }
}

__device__ __forceinline__ uint32_t shuffle_out_lut(uint32_t w, const uint32_t (*tab)[256]) {
    return tab[0][w & 255u] | tab[1][(w >> 8) & 255u] | tab[2][(w >> 16) & 255u] | tab[3][w >> 24];
}

__global__ void k_quantise_frames(SectorArgs a, const float *in, int64_t nframes, uint8_t *out) {
    __shared__ float lv[256];
    const int nl = a.out_nlevels;
    for (int i = threadIdx.x; i < nl && i < 256; i += blockDim.x) lv[i] = a.out_levels[i];
    __syncthreads();
    const int nwo = a.bpf_out / 4;
    const uint32_t mask = (a.bit_out >= 32) ? 0xFFFFFFFFu : ((1u << a.bit_out) - 1u);
    const int64_t f = blockIdx.x;
    const float *x = in + f * a.N;
    for (int w = threadIdx.x; w < nwo; w += blockDim.x) {
        uint32_t word = quant_word([x](int i) { return x[i]; }, a.N, w, a.bit_out, out_qlev(a, lv), mask);
        reinterpret_cast<uint32_t *>(out + f * a.bpf_out)[w] =
            a.out_shuffle_ident ? word : shuffle_out_word(word, a.out_shifts);
    }
}

// The inverse FFT and the quantiser fused for the phased .raw output: the
// time series never leaves LDS.  Its samples are the floats k_irfft_frames
// writes (z.x * scale, odd_sign * (z.y * scale)), so the words are the same.
// Each workgroup takes G frames at a time (T threads each; G = 2 from 128
// threads per frame, else 256 threads in all) and walks frame groups
// blockIdx.x, blockIdx.x + gridDim.x, ...: its levels, output shuffle tables
// (dynamic LDS, only for a shuffled output) and, for M <= 2048, FFT
// twiddles are staged once per workgroup.
template <int T>
constexpr int irq_frames() { return T >= 128 ? 2 : 256 / T; }

template <int E, int T>
__global__ __launch_bounds__(irq_frames<T>() * T) void k_irfft_quantise(SectorArgs a, const float2 *in,
                                                          int64_t nframes, uint8_t *out,
                                                          int negate_odd) {
    constexpr int M = E * T;
    constexpr int MP = M + M / 16 + 2;
    constexpr bool TWL = M <= 2048;   // FFT twiddles in LDS (at most 16 KB)
    constexpr int G = irq_frames<T>();   // frames per workgroup pass
    constexpr int NT = G * T;
    __shared__ float2 zbuf[G][MP];
    __shared__ float2 twl[TWL ? M : 1];
    extern __shared__ uint32_t dyn[];  // levels (256 floats), then the output shuffle tables
    float *lv = reinterpret_cast<float *>(dyn);
    uint32_t(*oshuf)[256] = reinterpret_cast<uint32_t(*)[256]>(dyn + 256);
    const int s = T >= 64 ? __builtin_amdgcn_readfirstlane((int)threadIdx.x / T) : (int)threadIdx.x / T;
    const int t = threadIdx.x - s * T;
    const int nl = a.out_nlevels < 256 ? a.out_nlevels : 256;
    for (int i = threadIdx.x; i < nl; i += NT) lv[i] = a.out_levels[i];
    const bool ident = a.out_shuffle_ident != 0;
    if (!ident) build_out_shuffle(oshuf, a.out_shifts, threadIdx.x, NT);
    if constexpr (TWL)
        for (int i = threadIdx.x; i < M; i += NT) twl[i] = a.tw_fft[i];
    const float2 *tw = TWL ? twl : a.tw_fft;
    const float scale = 1.0f / (float)(2 * M);
    const float odd_sign = negate_odd ? 1.0f : -1.0f;   // main.rs:2290-2292
    const float2 *zs = zbuf[s];
    const int nwo = a.bpf_out / 4;
    const uint32_t mask = (a.bit_out >= 32) ? 0xFFFFFFFFu : ((1u << a.bit_out) - 1u);
    const QLev ql = out_qlev(a, lv);
    auto sample = [zs, scale, odd_sign](int i) {
        const float2 z = zs[pad(i >> 1)];
        return (i & 1) ? odd_sign * (z.y * scale) : z.x * scale;
    };
    const int64_t ngroups = (nframes + G - 1) / G;
    for (int64_t fp = blockIdx.x; fp < ngroups; fp += gridDim.x) {
        const int64_t f = G * fp + s;
        const bool live = f < nframes;
        const int tl = loop_local(t);   // keeps per-thread addresses out of loop-invariant registers
        float2 v[E];
        if (live) irfft_prepare<E, T>(in + f * (M + 1), M, a.tw_post, v, tl);
        else
            for (int r = 0; r < E; ++r) v[r] = make_float2(0.f, 0.f);
        fft_forward<E, T>(zbuf[s], v, tw, tl);   // ends with a barrier: zbuf and the tables complete
        if (live)
            for (int w = tl; w < nwo; w += T) {
                const uint32_t word = quant_word_any(sample, 2 * M, w, a.bit_out, ql, mask);
                reinterpret_cast<uint32_t *>(out + f * a.bpf_out)[w] =
                    ident ? word : shuffle_out_lut(word, oshuf);
            }
        __syncthreads();   // zbuf is rewritten by the next group's FFT
    }
}

// The second half of the inverse real FFT at N = 16384 (M = 128 x 64) and
// the quantiser.  Input per frame: the 128 x 64 block the phased second pass
// wrote, B[k1][n2] = W_M^{n2 k1} DFT-64_{k2}(y[k1 + 128 k2]) with y the
// conjugated half-length spectrum.  Column n2's DFT-128 over k1 gives the
// transform of y at n2 + 64 n1, split as k1 = j + 8 r, n1 = a + 16 b:
// thread (n2, j) takes the DFT-16 over r, times W_128^{j a}; after an LDS
// transpose thread (n2, a pair) takes the DFT-8 over j for two a, writing
// the natural-order series to LDS; then the samples and words as
// k_irfft_quantise (x[2 n] = Re, x[2 n + 1] = odd_sign Im, times 1 / N).
// One frame per workgroup pass (512 threads), frames blockIdx.x, +
// gridDim.x, ...
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void k_irfft_cols_quantise(SectorArgs a, const float2 *in,
                                                             int64_t nframes, uint8_t *out,
                                                             int negate_odd, const float2 *tw_m) {
    constexpr int M = 8192, M2 = 64, ZP = M + M / 16 + 2;
    __shared__ float2 zb[ZP];
    __shared__ float2 w128[128];
    extern __shared__ uint32_t dyn[];  // levels (256 floats), then the output shuffle tables
    float *lv = reinterpret_cast<float *>(dyn);
    uint32_t(*oshuf)[256] = reinterpret_cast<uint32_t(*)[256]>(dyn + 256);
    const int t0 = threadIdx.x, t = t0;
    const int jw = __builtin_amdgcn_readfirstlane(t >> 6);   // j (first phase), a pair (second)
    const int nl = a.out_nlevels < 256 ? a.out_nlevels : 256;
    for (int i = t; i < nl; i += 512) lv[i] = a.out_levels[i];
    const bool ident = a.out_shuffle_ident != 0;
    if (!ident) build_out_shuffle(oshuf, a.out_shifts, t, 512);
    if (t < 128) w128[t] = tw_m[64 * t];
    __syncthreads();
    const float scale = 1.0f / (float)(2 * M);
    const float odd_sign = negate_odd ? 1.0f : -1.0f;   // main.rs:2290-2292
    const int nwo = a.bpf_out / 4;
    const uint32_t mask = (a.bit_out >= 32) ? 0xFFFFFFFFu : ((1u << a.bit_out) - 1u);
    const QLev ql = out_qlev(a, lv);
    auto sample = [scale, odd_sign](int i) {
        const float2 z = zb[pad(i >> 1)];
        return (i & 1) ? odd_sign * (z.y * scale) : z.x * scale;
    };
    for (int64_t f = blockIdx.x; f < nframes; f += gridDim.x) {
        const int n2 = loop_local(t0) & 63;   // keeps per-thread addresses out of loop-invariant registers
        const float2 *B = in + f * M;
        float2 v[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) v[r] = B[(jw + 8 * r) * M2 + n2];
        dftR<16>(v);                                   // over r: index a
#pragma unroll
        for (int q = 1; q < 16; ++q) v[q] = c_mul(v[q], w128[jw * q]);
#pragma unroll
        for (int q = 0; q < 16; ++q) zb[pad((q * 8 + jw) * 64 + n2)] = v[q];
        __syncthreads();
        float2 u[2][8];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int q = 0; q < 8; ++q) u[h][q] = zb[pad(((2 * jw + h) * 8 + q) * 64 + n2)];
        __syncthreads();
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            dftR<8>(u[h]);                             // over j: index b
#pragma unroll
            for (int b = 0; b < 8; ++b) zb[pad(n2 + 64 * (2 * jw + h + 16 * b))] = u[h][b];
        }
        __syncthreads();
        for (int w = t; w < nwo; w += 512) {
            const uint32_t word = quant_word_any(sample, 2 * M, w, a.bit_out, ql, mask);
            reinterpret_cast<uint32_t *>(out + f * a.bpf_out)[w] = ident ? word : shuffle_out_lut(word, oshuf);
        }
        __syncthreads();   // zb is rewritten by the next frame
    }
}

hipError_t launch_irfft_cols_quantise(const SectorArgs &a, const float2 *in, int64_t nframes,
                                      uint8_t *out, bool negate_odd, const float2 *tw_m,
                                      hipStream_t s) {
    if (nframes <= 0) return hipSuccess;
    if (a.N != 16384 || a.M != 8192) return hipErrorInvalidValue;
    const size_t dyn = (size_t)(256 + (a.out_shuffle_ident ? 0 : 4 * 256)) * 4;
    const int64_t grid = std::min<int64_t>(nframes, 2048);
    hipLaunchKernelGGL(k_irfft_cols_quantise, dim3((unsigned)grid), dim3(512), dyn, s, a, in, nframes,
                       out, negate_odd ? 1 : 0, tw_m);
    return hipGetLastError();
}

template <int E, int T>
static hipError_t launch_irfft_quantise_t(const SectorArgs &a, const float2 *in, int64_t nframes,
                                          uint8_t *out, bool negate_odd, hipStream_t s) {
    // at most 2^22 threads (several resident rounds on 256 CUs); each
    // workgroup walks groups of G frames
    constexpr int G = irq_frames<T>();
    const int64_t ngroups = (nframes + G - 1) / G;
    const int64_t cap = ((int64_t)1 << 22) / (G * T);
    const size_t dyn = (size_t)(256 + (a.out_shuffle_ident ? 0 : 4 * 256)) * 4;
    hipLaunchKernelGGL((k_irfft_quantise<E, T>), dim3((unsigned)std::min<int64_t>(ngroups, cap)),
                       dim3(G * T), dyn, s, a, in, nframes, out, negate_odd ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_irfft_quantise(const SectorArgs &a, const float2 *in, int64_t nframes,
                                 uint8_t *out, bool negate_odd, hipStream_t s) {
    if (nframes <= 0) return hipSuccess;
    switch (a.N) {
    case 64: return launch_irfft_quantise_t<2, 16>(a, in, nframes, out, negate_odd, s);
    case 128: return launch_irfft_quantise_t<2, 32>(a, in, nframes, out, negate_odd, s);
    case 256: return launch_irfft_quantise_t<4, 32>(a, in, nframes, out, negate_odd, s);
    case 512: return launch_irfft_quantise_t<4, 64>(a, in, nframes, out, negate_odd, s);
    case 1024: return launch_irfft_quantise_t<8, 64>(a, in, nframes, out, negate_odd, s);
    case 2048: return launch_irfft_quantise_t<8, 128>(a, in, nframes, out, negate_odd, s);
    case 4096: return launch_irfft_quantise_t<8, 256>(a, in, nframes, out, negate_odd, s);
    case 8192: return launch_irfft_quantise_t<8, 512>(a, in, nframes, out, negate_odd, s);
    case 16384: return launch_irfft_quantise_t<16, 512>(a, in, nframes, out, negate_odd, s);
    default: return hipErrorInvalidValue;
    }
}

// The quantiser reading the inverse FFT's complex output z (M = N/2 points
// per frame, natural order) of the HBM four-step path: sample 2m is
// z[m].x * scale, sample 2m + 1 is odd_sign * (z[m].y * scale), the floats
// the staged finish kernel writes, so the words match.
// One word per thread over a (frame, 256-word tile) grid: a frame of
// N = 65536 samples has up to 16,384 words, and a chunk's few hundred frames
// alone would leave the GPU at about two waves per SIMD.
__global__ __launch_bounds__(256) void k_quantise_cplx(SectorArgs a, const float2 *z,
                                                       int64_t nframes, int negate_odd,
                                                       uint8_t *out) {
    __shared__ float lv[256];
    const int nl = a.out_nlevels < 256 ? a.out_nlevels : 256;
    for (int i = threadIdx.x; i < nl; i += blockDim.x) lv[i] = a.out_levels[i];
    __syncthreads();
    const int nwo = a.bpf_out / 4;
    const int w = (int)blockIdx.y * (int)blockDim.x + (int)threadIdx.x;
    if (w >= nwo) return;
    const uint32_t mask = (a.bit_out >= 32) ? 0xFFFFFFFFu : ((1u << a.bit_out) - 1u);
    const int64_t f = blockIdx.x;
    const float2 *zf = z + f * (a.N / 2);
    const float scale = 1.0f / (float)a.N;
    const float odd_sign = negate_odd ? 1.0f : -1.0f;   // main.rs:2290-2292
    const uint32_t word = quant_word(
        [zf, scale, odd_sign](int i) {
            const float2 v = zf[i >> 1];
            return (i & 1) ? odd_sign * (v.y * scale) : v.x * scale;
        },
        a.N, w, a.bit_out, out_qlev(a, lv), mask);
    reinterpret_cast<uint32_t *>(out + f * a.bpf_out)[w] =
        a.out_shuffle_ident ? word : shuffle_out_word(word, a.out_shifts);
}

hipError_t launch_quantise_cplx(const SectorArgs &a, const float2 *z, int64_t nframes,
                                bool negate_odd, uint8_t *out, hipStream_t s) {
    if (nframes <= 0) return hipSuccess;
    const int nwo = a.bpf_out / 4;
    hipLaunchKernelGGL(k_quantise_cplx, dim3((unsigned)nframes, (unsigned)((nwo + 255) / 256)), dim3(256),
                       0, s, a, z, nframes, negate_odd ? 1 : 0, out);
    return hipGetLastError();
}

hipError_t launch_quantise_frames(const SectorArgs &a, const float *in, int64_t nframes,
                                  uint8_t *out, hipStream_t s) {
    if (nframes <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_quantise_frames, dim3((unsigned)nframes), dim3(256), 0, s, a, in, nframes,
                       out);
    return hipGetLastError();
}

}  // namespace fxc
