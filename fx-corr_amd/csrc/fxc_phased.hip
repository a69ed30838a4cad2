// Phased-array sum of the two stations (yi-phasedarray; main.rs:2222-2302,
// 2391-2411): per frame, decode + integer shift + real FFT of both stations
// (the fused sector kernel's front end), each station's fractional-delay
// rotation (utils.rs:69-124 with rate = accel = t = 0, USB sign), fringe
// phasors, band alignment onto the output grid, cb = w1 s1 + w2 s2, |cb|^2
// accumulated in FP64 and, for the .raw output, cb with DC/Nyquist imaginary
// parts zeroed written out for the inverse FFT + quantise kernels.
// PREC 0 (plain, reference-faithful): the f32 rotation chain itself comes
// from k_phase_chain, bit for bit, and the fringe phasors are the
// reference's f32 values.  PREC 1 (chosen): bin b's rotation exp(2 pi i b th)
// (th = -df frac, turns per bin) from the FP64 phase, reduced mod 1, with
// f32 sincos for a thread's first bin and FP32 steps after it, and fringe
// phasors from the FP64 phase -lo tau in turns; no chain table.
#include "fxc_dev.h"

namespace fxc {

template <int PREC>
__device__ void make_phased_param(int64_t idx, const SectorArgs &a, PhasedParam &p) {
    FrameDelayDev d;
    compute_frame_delay(idx, a.dly, d, nullptr, nullptr);
    p.s1 = clamp_shift(d.int1, a.N);
    p.s2 = clamp_shift(d.int2, a.N);
    if constexpr (PREC == 0) {
        p.fr1 = make_float2(d.fr1[0], d.fr1[1]);
        p.fr2 = make_float2(d.fr2[0], d.fr2[1]);
        p.th1 = p.th2 = 0.0;
    } else {
        double tau1, tau2, tf1, tf2;
        frame_taus(idx, a.dly, tau1, tau2, tf1, tf2);
        const float w1 = a.dly.lo1 == 0.0 ? 0.0f : (float)wrap_turns(-a.dly.lo1 * tf1);
        const float w2 = a.dly.lo2 == 0.0 ? 0.0f : (float)wrap_turns(-a.dly.lo2 * tf2);
        p.fr1 = make_float2(__builtin_amdgcn_cosf(w1), __builtin_amdgcn_sinf(w1));
        p.fr2 = make_float2(__builtin_amdgcn_cosf(w2), __builtin_amdgcn_sinf(w2));
        p.th1 = -a.df * d.frac1;
        p.th2 = -a.df * d.frac2;
    }
}

// accumulate = 0: the workgroup's partial slot is set; 1: added to (the
// sector is processed in chunks that reuse the same slots in order).
// MF (chosen form, levels exact in FP16): the FFT's first pass (radix 32)
// as FP16 hi/lo MFMA GEMMs with the operand table in LDS, as k_sector's.
template <int E, int T, int PREC, bool MF = false>
__global__ __launch_bounds__(2 * T) void k_phased(SectorArgs a, int accumulate) {
    constexpr int M = E * T;
    constexpr int MP = M + M / 16 + 2;
    constexpr int NT = 2 * T;
    constexpr int NB = (M + NT) / NT;
    // chain prefetch: fits in registers (the one-wave N = 256 form spills)
    constexpr bool PF = PREC == 0 && NB <= 5 && T >= 64;
    __shared__ float2 zbuf[2][MP];
    __shared__ PhasedParam pp[kFrameBatch];
    __shared__ float lev[2][256];
    __shared__ uint4 b1s[MF ? 16 * 64 : 1];

    const int tid0 = threadIdx.x;
    const int s = half_index<T>(tid0);
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
    if constexpr (MF)
        for (int i = tid0; i < 16 * 64; i += NT) b1s[i] = reinterpret_cast<const uint4 *>(a.mf_tab)[i];

    double ph[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) ph[i] = 0.0;

    const int gsn = a.grid;                          // station on the output grid
    const int gs = (gsn == 0) ? a.a1s : a.a2s;
    const int ge = gs + a.ov;
    const int sb_other = (gsn == 0) ? (a.a2s - a.a1s) : (a.a1s - a.a2s);
    const float w1 = (float)a.w1, w2 = (float)a.w2;
    const bool write_cb = (a.flags & FLAG_RAW_OUT) != 0 && a.cb_out != nullptr;
    float2 *zs = zbuf[s];
    uint32_t *ws = reinterpret_cast<uint32_t *>(zbuf[s]);
    const float2 *zg = zbuf[gsn];
    const float2 *zo = zbuf[1 - gsn];
    __syncthreads();

    for (int64_t fb = f_begin; fb < f_end; fb += kFrameBatch) {
        const int nb = (int)((f_end - fb) < kFrameBatch ? (f_end - fb) : kFrameBatch);
        if (tid < nb) make_phased_param<PREC>(a.first_frame + fb + tid, a, pp[tid]);
        __syncthreads();
        for (int jf = 0; jf < nb; ++jf) {
            const int64_t f = fb + jf;
            const int tid = loop_local(tid0), t = loop_local(t0);   // see loop_local
            // rotation of bin b: the reference's own f32 chain r_b
            // (k_phase_chain, rows 2f + station of this chunk).  Loaded
            // before the decode and FFT so their latency overlaps them.
            const float2 *chg = PREC ? nullptr : a.chain + (2 * f + gsn) * a.chain_ld;
            const float2 *cho = PREC ? nullptr : a.chain + (2 * f + 1 - gsn) * a.chain_ld;
            float2 rg[PF ? NB : 1], ro[PF ? NB : 1];
            if constexpr (PF) {
#pragma unroll
                for (int i = 0; i < NB; ++i) {
                    const int g = tid + NT * i;
                    rg[i] = g <= M ? chg[g] : make_float2(0.f, 0.f);
                    ro[i] = (g <= M && g >= gs && g < ge) ? cho[g + sb_other] : make_float2(0.f, 0.f);
                }
            }
            stage_words<T>(raw + f * st.bpf, st, ws, t);
            __syncthreads();
            const int sh = s ? pp[jf].s2 : pp[jf].s1;
            float2 v[E];
            if constexpr (MF) {
                // each thread decodes a run of E consecutive elements into LDS
                decode_run<E>(ws, st, lev[s], t * E, sh, N, v);
                __syncthreads();
#pragma unroll
                for (int r = 0; r < E; ++r) zs[pad(t * E + r)] = v[r];
                __syncthreads();
                stockham_first_mfma<E, T>(zs, b1s, t);
                stockham_rest<E, T, 32>(zs, a.tw_fft, t);
            } else {
#pragma unroll
                for (int r = 0; r < E; ++r) v[r] = decode_elem(ws, st, lev[s], t + T * r, sh, N);
                __syncthreads();
                fft_forward<E, T>(zs, v, a.tw_fft, t);
            }

            const PhasedParam &P = pp[jf];
            const float2 frg = gsn == 0 ? P.fr1 : P.fr2;
            const float2 fro = gsn == 0 ? P.fr2 : P.fr1;
            // chosen: the thread's rotations, bin tid + NT i = first * step^i
            float2 Rg = make_float2(1.f, 0.f), Ro = Rg, Sg = Rg, So = Rg;
            if constexpr (PREC != 0) {
                const double thg = gsn == 0 ? P.th1 : P.th2, tho = gsn == 0 ? P.th2 : P.th1;
                Rg = phasor_turns((double)tid * thg);
                Ro = phasor_turns((double)(tid + sb_other) * tho);
                Sg = phasor_turns((double)NT * thg);
                So = phasor_turns((double)NT * tho);
            }
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                const int g = tid + NT * i;
                if (PREC != 0 && i > 0) {
                    Rg = c_mul(Rg, Sg);
                    Ro = c_mul(Ro, So);
                }
                if (g <= M) {
                    float2 sg = c_mul(untangle(zg, g, M, a.tw_post),
                                      PREC ? Rg : (PF ? rg[PF ? i : 0] : chg[g]));
                    if (g == M) sg.y = 0.0f;                  // Nyquist (utils.rs:118-123)
                    const float2 cg = c_mul(sg, frg);
                    float2 co = make_float2(0.0f, 0.0f);
                    if (g >= gs && g < ge) {
                        const int j = g + sb_other;
                        float2 so = c_mul(untangle(zo, j, M, a.tw_post),
                                          PREC ? Ro : (PF ? ro[PF ? i : 0] : cho[j]));
                        if (j == M) so.y = 0.0f;
                        co = c_mul(so, fro);
                    }
                    const float2 c1 = gsn == 0 ? cg : co, c2 = gsn == 0 ? co : cg;
                    float2 cb;
                    {
#pragma clang fp contract(off)
                        const float2 a1 = c_scale(c1, w1), a2 = c_scale(c2, w2);
                        cb = make_float2(a1.x + a2.x, a1.y + a2.y);
                    }
                    ph[i] += (double)norm_sqr(cb);
                    if (write_cb) {
                        if (g == 0 || g == M) cb.y = 0.0f;     // main.rs:2283-2284
                        a.cb_out[f * (M + 1) + g] = cb;
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
            if (accumulate) {
                out[P_PH * half + g] += ph[i];
            } else {   // the other products stay zero over the later chunks
                out[P_PH * half + g] = ph[i];
                out[P_11 * half + g] = 0.0;
                out[P_22 * half + g] = 0.0;
                out[P_12R * half + g] = 0.0;
                out[P_12I * half + g] = 0.0;
            }
        }
    }
}

template <int E, int T>
static hipError_t launch_phased_t(const SectorArgs &a, int precision, int accumulate, int n_wg,
                                  hipStream_t s) {
    if (precision == 0) {
        hipLaunchKernelGGL((k_phased<E, T, 0>), dim3(n_wg), dim3(2 * T), 0, s, a, accumulate);
        return hipGetLastError();
    }
    // a.mf_tab: the chosen plan's first FFT pass runs on FP16 MFMA (whole
    // waves of 32-point DFT rows: E >= 8, T a multiple of 64)
    if constexpr (E >= 8 && T % 64 == 0) {
        if (a.mf_tab != nullptr) {
            hipLaunchKernelGGL((k_phased<E, T, 1, true>), dim3(n_wg), dim3(2 * T), 0, s, a, accumulate);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((k_phased<E, T, 1>), dim3(n_wg), dim3(2 * T), 0, s, a, accumulate);
    return hipGetLastError();
}

hipError_t launch_phased(const SectorArgs &a, int precision, int accumulate, int n_wg,
                         hipStream_t s) {
    if (n_wg <= 0) return hipSuccess;
    switch (a.N) {
    case 64: return launch_phased_t<2, 16>(a, precision, accumulate, n_wg, s);
    case 128: return launch_phased_t<2, 32>(a, precision, accumulate, n_wg, s);
    case 256: return launch_phased_t<4, 32>(a, precision, accumulate, n_wg, s);
    case 512: return launch_phased_t<4, 64>(a, precision, accumulate, n_wg, s);
    case 1024: return launch_phased_t<8, 64>(a, precision, accumulate, n_wg, s);
    case 2048: return launch_phased_t<8, 128>(a, precision, accumulate, n_wg, s);
    case 4096: return launch_phased_t<8, 256>(a, precision, accumulate, n_wg, s);
    case 8192: return launch_phased_t<8, 512>(a, precision, accumulate, n_wg, s);
    case 16384: return launch_phased_t<16, 512>(a, precision, accumulate, n_wg, s);
    default: return hipErrorInvalidValue;
    }
}

// ------------------------------------------------ N = 32768 / 65536
// The same sum through HBM scratch: per-frame parameters (and the decode's
// integer shifts), then, after the HBM four-step FFTs of both stations, one
// output-grid bin per thread over a group of frames (blockIdx.y), with the
// arithmetic of k_phased's epilogue; partial slot per frame group.
template <int PREC>
__global__ void k_phased_params(SectorArgs a, PhasedParam *pp, int64_t *shifts) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= a.nf) return;
    PhasedParam p;
    make_phased_param<PREC>(a.first_frame + f, a, p);
    pp[f] = p;
    shifts[2 * f] = p.s1;
    shifts[2 * f + 1] = p.s2;
}

template <int PREC>
__global__ __launch_bounds__(256) void k_phased_large(SectorArgs a, const float2 *z1, const float2 *z2,
                                                      const PhasedParam *pp, int64_t fpg,
                                                      int accumulate) {
    const int M = a.M;
    const int half = M + 1;
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t f0 = (int64_t)blockIdx.y * fpg;
    int64_t f1 = f0 + fpg;
    if (f1 > a.nf) f1 = a.nf;
    if (g >= half) return;
    const int gsn = a.grid;
    const int gs = (gsn == 0) ? a.a1s : a.a2s;
    const bool in_ov = g >= gs && g < gs + a.ov;
    const int j = g + ((gsn == 0) ? (a.a2s - a.a1s) : (a.a1s - a.a2s));   // other station's bin
    const float w1 = (float)a.w1, w2 = (float)a.w2;
    const bool write_cb = (a.flags & FLAG_RAW_OUT) != 0 && a.cb_out != nullptr;
    const float2 *zg = gsn == 0 ? z1 : z2;
    const float2 *zo = gsn == 0 ? z2 : z1;
    double ph = 0.0;
    for (int64_t f = f0; f < f1; ++f) {
        const PhasedParam P = pp[f];
        const float2 frg = gsn == 0 ? P.fr1 : P.fr2;
        const float2 fro = gsn == 0 ? P.fr2 : P.fr1;
        float2 rg, ro = make_float2(1.f, 0.f);
        if constexpr (PREC == 0) {
            rg = a.chain[(2 * f + gsn) * a.chain_ld + g];
            if (in_ov) ro = a.chain[(2 * f + 1 - gsn) * a.chain_ld + j];
        } else {
            rg = phasor_turns((double)g * (gsn == 0 ? P.th1 : P.th2));
            if (in_ov) ro = phasor_turns((double)j * (gsn == 0 ? P.th2 : P.th1));
        }
        float2 sg = c_mul(untangle_np(zg + f * M, g, M, a.tw_post), rg);
        if (g == M) sg.y = 0.0f;                          // Nyquist (utils.rs:118-123)
        const float2 cg = c_mul(sg, frg);
        float2 co = make_float2(0.0f, 0.0f);
        if (in_ov) {
            float2 so = c_mul(untangle_np(zo + f * M, j, M, a.tw_post), ro);
            if (j == M) so.y = 0.0f;
            co = c_mul(so, fro);
        }
        const float2 c1 = gsn == 0 ? cg : co, c2 = gsn == 0 ? co : cg;
        float2 cb;
        {
#pragma clang fp contract(off)
            const float2 a1 = c_scale(c1, w1), a2 = c_scale(c2, w2);
            cb = make_float2(a1.x + a2.x, a1.y + a2.y);
        }
        ph += (double)norm_sqr(cb);
        if (write_cb) {
            if (g == 0 || g == M) cb.y = 0.0f;            // main.rs:2283-2284
            a.cb_out[f * half + g] = cb;
        }
    }
    double *out = a.part + (size_t)blockIdx.y * kNumProducts * half;
    out[P_PH * half + g] = accumulate ? out[P_PH * half + g] + ph : ph;
    out[P_11 * half + g] = 0.0;
    out[P_22 * half + g] = 0.0;
    out[P_12R * half + g] = 0.0;
    out[P_12I * half + g] = 0.0;
}

size_t phased_param_bytes() { return sizeof(PhasedParam); }

hipError_t launch_phased_params(const SectorArgs &a, int precision, void *pp, int64_t *shifts,
                                hipStream_t s) {
    if (a.nf <= 0) return hipSuccess;
    if (precision == 0)
        hipLaunchKernelGGL(k_phased_params<0>, dim3((unsigned)((a.nf + 127) / 128)), dim3(128), 0, s,
                           a, reinterpret_cast<PhasedParam *>(pp), shifts);
    else
        hipLaunchKernelGGL(k_phased_params<1>, dim3((unsigned)((a.nf + 127) / 128)), dim3(128), 0, s,
                           a, reinterpret_cast<PhasedParam *>(pp), shifts);
    return hipGetLastError();
}

hipError_t launch_phased_large(const SectorArgs &a, int precision, const float2 *z1,
                               const float2 *z2, const void *pp, int n_groups, int accumulate,
                               hipStream_t s) {
    const int64_t fpg = (a.nf + n_groups - 1) / n_groups;
    const dim3 grid((unsigned)((a.M + 1 + 255) / 256), (unsigned)n_groups);
    if (precision == 0)
        hipLaunchKernelGGL(k_phased_large<0>, grid, dim3(256), 0, s, a, z1, z2,
                           reinterpret_cast<const PhasedParam *>(pp), fpg, accumulate);
    else
        hipLaunchKernelGGL(k_phased_large<1>, grid, dim3(256), 0, s, a, z1, z2,
                           reinterpret_cast<const PhasedParam *>(pp), fpg, accumulate);
    return hipGetLastError();
}

// dst[i] += src[i] (FP64), e.g. a job's spectra += one chunk's spectra.
__global__ void k_add_f64(double *dst, const double *src, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] += src[i];
}

hipError_t launch_add_f64(double *dst, const double *src, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_add_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dst, src, n);
    return hipGetLastError();
}

// Whole-scan cross spectrum: ac[2k], ac[2k+1] += s * (re[k], im[k]), k < half
// (main.rs:1850: ac[k] += bc[k] * inv_fft2).
__global__ void k_scale_add_xcf(double *ac, const double *re, const double *im, int half, double s) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < half) {
        ac[2 * k] += re[k] * s;
        ac[2 * k + 1] += im[k] * s;
    }
}

hipError_t launch_scale_add_xcf(double *ac, const double *re, const double *im, int half, double s,
                                hipStream_t st) {
    if (half <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scale_add_xcf, dim3((unsigned)((half + 255) / 256)), dim3(256), 0, st, ac, re,
                       im, half, s);
    return hipGetLastError();
}

// Graph-safe zero fill (a kernel, not a memset node).
__global__ void k_fill_zero(uint32_t *p, size_t nwords, uint8_t *tail, size_t ntail) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = i; k < nwords; k += stride) p[k] = 0u;
    for (size_t k = i; k < ntail; k += stride) tail[k] = 0;
}

hipError_t launch_fill_zero(void *p, size_t bytes, hipStream_t s) {
    if (bytes == 0) return hipSuccess;
    const size_t nwords = ((uintptr_t)p % 4 == 0) ? bytes / 4 : 0;
    uint8_t *tail = (uint8_t *)p + nwords * 4;
    const size_t ntail = bytes - nwords * 4;
    size_t blocks = (nwords + 255) / 256;
    if (blocks < (ntail + 255) / 256) blocks = (ntail + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)blocks), dim3(256), 0, s, (uint32_t *)p, nwords,
                       tail, ntail);
    return hipGetLastError();
}

}  // namespace fxc
