// FX path for FFT lengths beyond one workgroup's LDS (N = 32768, 65536:
// M = N/2 complex points = 128/256 KB of FP32 per station-frame).  Frames
// are processed in chunks through HBM scratch:
//
//   1. k_frame_params   per-frame delay entries (main.rs:184-241) and the
//                       XCF phase start/step (main.rs:133-149)
//   2. k_decode_frames  decode + integer shift to FP32 (the bit-exact decode)
//   3. k_fft_pass       four-step FFT, m = M2 m1 + m2, k = k1 + M1 k2:
//                       column DFTs over m1 (length M1) times W_M^{m2 k1},
//                       then row DFTs over m2 (length M2) into Z in natural
//                       order; each pass is a batch of LDS Stockham FFTs
//   4. k_mac_large      realfft untangle and ACF/XCF products, one output
//                       bin per thread looping over a group of frames, FP64
//                       accumulators (any band shift)
#include "fxc_dev.h"

namespace fxc {

// ---------------------------------------------------------------- params
template <int PREC>
__global__ void k_frame_params(SectorArgs a, int logm, FrameParam *params, int64_t *shifts) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= a.nf) return;
    FrameParam p;
    make_frame_param<PREC>(a.first_frame + f, a, logm, p);
    params[f] = p;
    shifts[2 * f] = p.s1;
    shifts[2 * f + 1] = p.s2;
}

hipError_t launch_frame_params(const SectorArgs &a, int precision, FrameParam *params,
                               int64_t *shifts, hipStream_t s) {
    if (a.nf <= 0) return hipSuccess;
    const int logm = __builtin_ctz((unsigned)a.M) + 1;
    const dim3 grid((unsigned)((a.nf + 127) / 128));
    if (precision == 0)
        hipLaunchKernelGGL(k_frame_params<0>, grid, dim3(128), 0, s, a, logm, params, shifts);
    else
        hipLaunchKernelGGL(k_frame_params<1>, grid, dim3(128), 0, s, a, logm, params, shifts);
    return hipGetLastError();
}

// ------------------------------------------------------- four-step passes
// A batch of CB independent L-point FFTs (L = E*T), column c of the batch
// at in[c*ics + m*ims] (m < L), result k written to out[c*ocs + k*oks],
// multiplied by W_M^{(c0 + c) k} when `twiddle` (tw_m = W_M^j, j < M).
// Loads and stores walk the unit-stride index across the threads.
template <int E, int T, int CB>
__global__ __launch_bounds__(CB *T) void k_fft_pass(const float2 *in, float2 *out, int64_t nbatch,
                                                    int64_t cols, int64_t ics, int64_t ims,
                                                    int64_t ocs, int64_t oks, int64_t frame_in,
                                                    int64_t frame_out, const float2 *tw_l,
                                                    const float2 *tw_m, int M, int twiddle) {
    constexpr int L = E * T;
    constexpr int LP = L + L / 16 + 2;
    __shared__ float2 buf[CB][LP];
    const int tid = threadIdx.x;
    const int64_t blocks_per_frame = cols / CB;
    const int64_t f = blockIdx.x / blocks_per_frame;
    const int64_t c0 = (blockIdx.x % blocks_per_frame) * CB;
    if (f >= nbatch) return;
    const float2 *src = in + f * frame_in;
    float2 *dst = out + f * frame_out;
    // load: unit stride along c (ics == 1) or along m (ims == 1)
    for (int i = tid; i < CB * L; i += CB * T) {
        int c, m;
        if (ics == 1) { c = i % CB; m = i / CB; }
        else { m = i % L; c = i / L; }
        buf[c][pad(m)] = src[(c0 + c) * ics + (int64_t)m * ims];
    }
    __syncthreads();
    const int c = tid / T, t = tid % T;
    float2 v[E];
#pragma unroll
    for (int r = 0; r < E; ++r) v[r] = buf[c][pad(t + T * r)];
    __syncthreads();
    fft_forward<E, T>(buf[c], v, tw_l, t);
    for (int i = tid; i < CB * L; i += CB * T) {
        int cc, k;
        if (ocs == 1) { cc = i % CB; k = i / CB; }
        else { k = i % L; cc = i / L; }
        float2 y = buf[cc][pad(k)];
        if (twiddle) y = c_mul(y, tw_m[((c0 + cc) * (int64_t)k) % M]);
        dst[(c0 + cc) * ocs + (int64_t)k * oks] = y;
    }
}

template <int E, int T>
static hipError_t launch_pass_t(const float2 *in, float2 *out, int64_t nbatch, int64_t cols,
                                int64_t ics, int64_t ims, int64_t ocs, int64_t oks,
                                int64_t frame_in, int64_t frame_out, const float2 *tw_l,
                                const float2 *tw_m, int M, int twiddle, hipStream_t s) {
    constexpr int CB = 16;
    if (cols % CB != 0) return hipErrorInvalidValue;
    const int64_t grid = nbatch * (cols / CB);
    hipLaunchKernelGGL((k_fft_pass<E, T, CB>), dim3((unsigned)grid), dim3(CB * T), 0, s, in, out,
                       nbatch, cols, ics, ims, ocs, oks, frame_in, frame_out, tw_l, tw_m, M,
                       twiddle);
    return hipGetLastError();
}

static hipError_t launch_pass(int L, const float2 *in, float2 *out, int64_t nbatch, int64_t cols,
                              int64_t ics, int64_t ims, int64_t ocs, int64_t oks,
                              int64_t frame_in, int64_t frame_out, const float2 *tw_l,
                              const float2 *tw_m, int M, int twiddle, hipStream_t s) {
    switch (L) {
    case 128: return launch_pass_t<4, 32>(in, out, nbatch, cols, ics, ims, ocs, oks, frame_in,
                                          frame_out, tw_l, tw_m, M, twiddle, s);
    case 256: return launch_pass_t<8, 32>(in, out, nbatch, cols, ics, ims, ocs, oks, frame_in,
                                          frame_out, tw_l, tw_m, M, twiddle, s);
    default: return hipErrorInvalidValue;
    }
}

// Z (natural order, M complex per station-frame) of nframes frames from the
// decoded samples x (N floats per frame): x read as z[m] = x[2m] + i x[2m+1].
// With mf_b1 the first pass's DFT-32 runs on FP16 MFMA (fxc_mfma.hip).
hipError_t launch_fft_large(const void *x, bool x_fp16, float2 *ybuf, float2 *z, int64_t nframes,
                            int M1, int M2, const float2 *tw1, const float2 *tw2,
                            const float2 *twm, const void *mf_b1, hipStream_t s) {
    const int M = M1 * M2;
    if (x_fp16 && !mf_b1) return hipErrorInvalidValue;   // FP16 input feeds the MFMA pass only
    const float2 *zin = reinterpret_cast<const float2 *>(x);
    // columns c = m2 (M2 of them): input z[m2 + M2 m1], output Y[k1][m2] * W_M^{m2 k1}
    hipError_t e = mf_b1 ? launch_fft_pass1_mfma(x, x_fp16, ybuf, nframes, M1, M2, mf_b1, twm, s)
                         : launch_pass(M1, zin, ybuf, nframes, M2, 1, M2, 1, M2, M, M, tw1, twm, M, 1, s);
    if (e != hipSuccess) return e;
    // rows c = k1 (M1 of them): input Y[k1][m2], output Z[k1 + M1 k2]
    return launch_pass(M2, ybuf, z, nframes, M1, M2, 1, 1, M1, M, M, tw2, twm, M, 0, s);
}

size_t frame_param_bytes() { return sizeof(FrameParam); }

__global__ void k_untangle_large(const float2 *z, int64_t nframes, int M, const float2 *tw_post,
                                 float2 *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t f = i / (M + 1);
    const int k = (int)(i % (M + 1));
    if (f >= nframes) return;
    out[i] = untangle_np(z + f * M, k, M, tw_post);
}

hipError_t launch_untangle_large(const float2 *z, int64_t nframes, int M, const float2 *tw_post,
                                 float2 *out, hipStream_t s) {
    const int64_t n = nframes * (M + 1);
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_untangle_large, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, z,
                       nframes, M, tw_post, out);
    return hipGetLastError();
}

// ------------------------------------------------ inverse real FFT, N >= 32768
// inverse_c2r_process (utils.rs:60-66) through the HBM four-step: conj(Z')
// of each frame's half spectrum, a forward FFT, then the real samples
// interleaved and scaled 1/N (odd samples negated once more for an LSB
// output stream, main.rs:2290-2292), as k_irfft_frames does in LDS.
__global__ void k_irfft_prepare_large(const float2 *X, int64_t nframes, int M, const float2 *tw_post,
                                      float2 *zc) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t f = i / M;
    const int k = (int)(i - f * M);
    if (f >= nframes) return;
    zc[i] = irfft_pre_elem(X + f * (M + 1), k, M, tw_post);
}

__global__ void k_irfft_finish_large(const float2 *z, int64_t nframes, int M, int negate_odd,
                                     float *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nframes * M) return;
    const float scale = 1.0f / (float)(2 * M);
    const float odd_sign = negate_odd ? 1.0f : -1.0f;
    const float2 v = z[i];
    out[2 * i] = v.x * scale;
    out[2 * i + 1] = odd_sign * (v.y * scale);
}

hipError_t launch_irfft_large(const float2 *X, int64_t nframes, int M1, int M2, const float2 *tw1,
                              const float2 *tw2, const float2 *twm, const float2 *tw_post,
                              float2 *zc, float2 *ybuf, float2 *z, bool negate_odd, float *out,
                              hipStream_t s) {
    const int M = M1 * M2;
    const int64_t n = nframes * M;
    if (n <= 0) return hipSuccess;
    const dim3 grid((unsigned)((n + 255) / 256));
    hipLaunchKernelGGL(k_irfft_prepare_large, grid, dim3(256), 0, s, X, nframes, M, tw_post, zc);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = launch_fft_large(zc, false, ybuf, z, nframes, M1, M2, tw1, tw2, twm, nullptr, s);
    if (e != hipSuccess || out == nullptr) return e;   // no out: the caller reads z
    hipLaunchKernelGGL(k_irfft_finish_large, grid, dim3(256), 0, s, z, nframes, M, negate_odd ? 1 : 0,
                       out);
    return hipGetLastError();
}

// ---------------------------------------------------------- untangle + MAC
// Bin g of the output grid per thread; frames [g0, g1) of the chunk; the
// per-frame phase terms come from k_frame_params.
template <int PREC>
__global__ __launch_bounds__(256) void k_mac_large(SectorArgs a, const float2 *z1, const float2 *z2,
                                                   const FrameParam *params, int64_t fpg,
                                                   int accumulate) {
    const int M = a.M;
    const int half = M + 1;
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t f0 = (int64_t)blockIdx.y * fpg;
    int64_t f1 = f0 + fpg;
    if (f1 > a.nf) f1 = a.nf;
    if (g >= half) return;
    const bool do_acf = (a.flags & FLAG_ACF) != 0;
    const bool do_xcf = (a.flags & FLAG_XCF) != 0;
    const bool ov_only = (a.flags & FLAG_ACF_OVERLAP) != 0;
    const int gs = (a.grid == 0) ? a.a1s : a.a2s;
    const bool in_ov = g >= gs && g < gs + a.ov;
    const int sb_other = (a.grid == 0) ? (a.a2s - a.a1s) : (a.a1s - a.a2s);
    const float2 *zg = a.grid == 0 ? z1 : z2;
    const float2 *zo = a.grid == 0 ? z2 : z1;
    const int kk = g - gs;   // overlap index (XCF phase exponent)
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t f = f0; f < f1; ++f) {
        const float2 Xg = untangle_np(zg + f * M, g, M, a.tw_post);
        float2 Xo = make_float2(0.0f, 0.0f);
        if (in_ov) Xo = untangle_np(zo + f * M, g + sb_other, M, a.tw_post);
        if (do_acf) {
            const float pgrid = (!ov_only || in_ov) ? norm_sqr(Xg) : 0.0f;
            const float poth = in_ov ? norm_sqr(Xo) : 0.0f;
            acc[0] += (double)(a.grid == 0 ? pgrid : poth);
            acc[1] += (double)(a.grid == 0 ? poth : pgrid);
        }
        if (in_ov && do_xcf) {
            const FrameParam &P = params[f];
            float2 phf;
            if (PREC == 0 && a.chain != nullptr) {
                phf = a.chain[f * a.chain_ld + kk];            // the reference's own chain
            } else if (PREC != 0) {
                // chosen form: the FP64 phase in turns, reduced mod 1, then the f32 sincos
                const float r = (float)wrap_turns(fma((double)kk, P.dth, P.phw));
                phf = make_float2(__builtin_amdgcn_cosf(r), __builtin_amdgcn_sinf(r));
            } else {
                cd ph = P.start;
                for (int b = 0; b < kMaxPow; ++b)
                    if ((kk >> b) & 1) ph = d_mul(ph, P.pw[b]);
                phf = make_float2((float)ph.x, (float)ph.y);
            }
            const float2 X1 = (a.grid == 0) ? Xg : Xo;
            const float2 X2 = (a.grid == 0) ? Xo : Xg;
            float2 v2 = c_mul(X1, c_conj(X2));
            if (PREC == 0) v2 = c_mul(v2, P.frmix);
            v2 = c_mul(v2, phf);
            acc[2] += (double)v2.x;
            acc[3] += (double)v2.y;
        }
    }
    double *out = a.part + (size_t)blockIdx.y * kNumProducts * half;
    const double prev[4] = {accumulate ? out[P_11 * half + g] : 0.0,
                            accumulate ? out[P_22 * half + g] : 0.0,
                            accumulate ? out[P_12R * half + g] : 0.0,
                            accumulate ? out[P_12I * half + g] : 0.0};
    out[P_PH * half + g] = 0.0;
    out[P_11 * half + g] = prev[0] + acc[0];
    out[P_22 * half + g] = prev[1] + acc[1];
    out[P_12R * half + g] = prev[2] + acc[2];
    out[P_12I * half + g] = prev[3] + acc[3];
}

// Wave-synchronous Stockham FFT of L = 64 E points in LDS (one wave, no
// workgroup barrier): v[r] = z[lane + 64 r] on entry, natural order in z on
// return.  Radix E passes (E = 2 or 4), then a radix-2 pass when needed.
__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int L, int R, int NS>
__device__ __forceinline__ void wave_stockham_pass(float2 *z, const float2 *tw, int lane) {
    if constexpr (NS < L) {
        constexpr int RR = (L / NS < R) ? (L / NS) : R;
        constexpr int BPT = L / (64 * RR);   // butterflies per lane
        float2 a[BPT][RR];
#pragma unroll
        for (int u = 0; u < BPT; ++u) {
            const int j = lane + 64 * u;
#pragma unroll
            for (int r = 0; r < RR; ++r) a[u][r] = z[pad(j + r * (L / RR))];
            if constexpr (NS > 1) {
                const int jm = j % NS;
#pragma unroll
                for (int r = 1; r < RR; ++r) a[u][r] = c_mul(a[u][r], tw[r * jm * (L / (NS * RR))]);
            }
            dftR<RR>(a[u]);
        }
        wave_sync_lds();
#pragma unroll
        for (int u = 0; u < BPT; ++u) {
            const int j = lane + 64 * u;
            const int base = (j / NS) * NS * RR + (j % NS);
#pragma unroll
            for (int r = 0; r < RR; ++r) z[pad(base + r * NS)] = a[u][r];
        }
        wave_sync_lds();
        wave_stockham_pass<L, R, NS * RR>(z, tw, lane);
    }
}

// ------------------------------------ second pass fused with the products
// Chosen form, both stations' bins on the same rows of the four-step (the
// band shift a multiple of M1).  Z[k1 + M1 k2] is row k1's DFT over m2 of
// Y[k1][m2], and bin M - k lies in row M1 - k1 (row 0 and row M1/2 are their
// own mirrors), so the untangle of a bin needs only its row pair.
// Workgroup (pair p, frame group): rows ra = p, rb = M1 - p (p = 0: rows 0
// and M1/2) of both stations; the pair's post twiddles W_N^{-g} and the
// DFT-M2 twiddles are staged in LDS once.  Per frame, wave q loads row q of
// [grid ra, grid rb, other ra, other rb] (the next two frames' rows are
// loaded into registers while this one is worked on) and runs its DFT-M2 in LDS
// with wave-level syncs; after one workgroup barrier each thread takes bins
// g = row + M1 k2 of the pair (and bin M in pair 0): the split-method
// untangle of both stations, the ACF/XCF products with the XCF phasor from
// the frame's FP64 phase (k_mac_large's arithmetic) and FP64 sums.  The
// rows are double-buffered in LDS, so a frame needs only that one barrier:
// a wave rewrites a buffer two frames later, after every thread has passed
// the next frame's barrier and so finished this frame's bins.  The frame
// group's partial slot holds the pair's bins contiguously (coalesced
// stores); the reduce over the slots puts them back in bin order.
__device__ __forceinline__ float2 untangle_ab(float2 A, float2 B, int k, int M, float2 w) {
    if (k == 0 || k == M) return make_float2(k == 0 ? A.x + A.y : A.x - A.y, 0.0f);
    const float ex = 0.5f * (A.x + B.x), ey = 0.5f * (A.y - B.y);
    const float dx = 0.5f * (A.x - B.x), dy = 0.5f * (A.y + B.y);
    return make_float2(ex + (w.x * dy - w.y * dx), ey - (w.x * dx + w.y * dy));
}

template <int M2>
__global__ __launch_bounds__(256) void k_large_pass2_mac(SectorArgs a, const float2 *y1,
                                                         const float2 *y2, const FrameParam *params,
                                                         int M1, const float2 *tw_m2, int64_t fpg,
                                                         int accumulate) {
    constexpr int E = M2 / 64;
    constexpr int LP = M2 + M2 / 16 + 2;
    constexpr int BPT = (2 * M2) / 256;   // bins per thread (plus bin M for thread 0 of pair 0)
    __shared__ float2 zr[2][4][LP];       // [buffer][grid ra, grid rb, other ra, other rb]
    __shared__ float2 twm[M2];
    __shared__ float2 twp[2][M2];         // tw_post[row + M1 k2], rows ra / rb
    const int M = a.M;
    const int half = M + 1;
    const int tid = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const int p = blockIdx.x;
    const int ra = p, rb = p == 0 ? M1 / 2 : M1 - p;
    for (int i = tid; i < M2; i += 256) twm[i] = tw_m2[i];
    for (int i = tid; i < 2 * M2; i += 256)
        twp[i / M2][i % M2] = a.tw_post[(i < M2 ? ra : rb) + M1 * (i % M2)];
    const int64_t g0 = (int64_t)blockIdx.y * fpg;
    int64_t g1 = g0 + fpg;
    if (g1 > a.nf) g1 = a.nf;
    const bool do_acf = (a.flags & FLAG_ACF) != 0;
    const bool do_xcf = (a.flags & FLAG_XCF) != 0;
    const bool ov_only = (a.flags & FLAG_ACF_OVERLAP) != 0;
    const int gs = (a.grid == 0) ? a.a1s : a.a2s;
    const int sbo = (a.grid == 0) ? (a.a2s - a.a1s) : (a.a1s - a.a2s);
    const int jo = sbo / M1;   // the other station's k2 offset (sbo is a multiple of M1)
    // this wave's row: q = w
    const float2 *src_row = ((w & 2) ? (a.grid == 0 ? y2 : y1) : (a.grid == 0 ? y1 : y2)) +
                            (int64_t)((w & 1) ? rb : ra) * M2;
    const bool nyq = p == 0 && tid == 0;
    double acc[BPT + 1][4];
#pragma unroll
    for (int i = 0; i <= BPT; ++i)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[i][c] = 0.0;
    // rows of frames f (cur), f + 1 and f + 2 in flight
    float2 cur[E], nx1[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        cur[e] = g0 < g1 ? src_row[g0 * (int64_t)M + lane + 64 * e] : make_float2(0.f, 0.f);
        nx1[e] = g0 + 1 < g1 ? src_row[(g0 + 1) * (int64_t)M + lane + 64 * e] : make_float2(0.f, 0.f);
    }
    __syncthreads();   // twiddles staged
    int b = 0;
    for (int64_t f = g0; f < g1; ++f, b ^= 1) {
        float2 nx2[E];
        const bool more = f + 2 < g1;
#pragma unroll
        for (int e = 0; e < E; ++e)
            nx2[e] = more ? src_row[(f + 2) * (int64_t)M + lane + 64 * e] : make_float2(0.f, 0.f);
        float2 *z = zr[b][w];
        dftR<E>(cur);
#pragma unroll
        for (int r = 0; r < E; ++r) z[pad(lane * E + r)] = cur[r];
        wave_sync_lds();
        wave_stockham_pass<M2, E, E>(z, twm, lane);
        __syncthreads();   // the four rows of frame f are ready
        const double dth = params[f].dth, phw = params[f].phw;
        const float2(*zb)[LP] = zr[b];
#pragma unroll
        for (int i = 0; i <= BPT; ++i) {
            int g, row, k2;
            if (i < BPT) {
                const int j = tid + 256 * i;
                row = j < M2 ? ra : rb;
                k2 = j % M2;
                g = row + M1 * k2;
            } else {
                if (!nyq) continue;
                row = 0;
                k2 = M2;
                g = M;
            }
            const int bg = row == ra ? 0 : 1;
            // mirror M - g: row 0 -> row 0, k2' = M2 - k2; else row M1 - row, k2' = M2 - 1 - k2
            const int bm = (row == 0 ? row : M1 - row) == ra ? 0 : 1;
            const int km = row == 0 ? (M2 - k2) & (M2 - 1) : M2 - 1 - k2;
            const float2 Zg = zb[bg][pad(k2 & (M2 - 1))], Zm = zb[bm][pad(km)];
            const float2 Xg = untangle_ab(Zg, Zm, g, M, twp[bg][k2 & (M2 - 1)]);
            const bool in_ov = g >= gs && g < gs + a.ov;
            float2 Xo = make_float2(0.0f, 0.0f);
            if (in_ov) {
                const int h = g + sbo, kh = k2 + jo;   // same row, k2 shifted
                const int khm = row == 0 ? (M2 - kh) & (M2 - 1) : M2 - 1 - kh;
                const float2 Oh = zb[2 + bg][pad(kh & (M2 - 1))];
                const float2 Om = zb[2 + bm][pad(khm & (M2 - 1))];
                Xo = untangle_ab(Oh, Om, h, M, twp[bg][kh & (M2 - 1)]);
            }
            if (do_acf) {
                const float pgrid = (!ov_only || in_ov) ? norm_sqr(Xg) : 0.0f;
                const float poth = in_ov ? norm_sqr(Xo) : 0.0f;
                acc[i][0] += (double)(a.grid == 0 ? pgrid : poth);
                acc[i][1] += (double)(a.grid == 0 ? poth : pgrid);
            }
            if (in_ov && do_xcf) {
                const float r = (float)wrap_turns(fma((double)(g - gs), dth, phw));
                const float2 phf = make_float2(__builtin_amdgcn_cosf(r), __builtin_amdgcn_sinf(r));
                const float2 X1 = (a.grid == 0) ? Xg : Xo;
                const float2 X2 = (a.grid == 0) ? Xo : Xg;
                const float2 v2 = c_mul(c_mul(X1, c_conj(X2)), phf);
                acc[i][2] += (double)v2.x;
                acc[i][3] += (double)v2.y;
            }
        }
#pragma unroll
        for (int e = 0; e < E; ++e) {
            cur[e] = nx1[e];
            nx1[e] = nx2[e];
        }
    }
    // partial slot in pair-major order: index p 2 M2 + j (j < M2: row ra,
    // k2 = j; else row rb, k2 = j - M2), bin M last; launch_reduce_pairs
    // restores bin order
    double *out = a.part + (size_t)blockIdx.y * kNumProducts * half;
#pragma unroll
    for (int i = 0; i <= BPT; ++i) {
        int g;
        if (i < BPT) {
            g = p * 2 * M2 + tid + 256 * i;
        } else {
            if (!nyq) continue;
            g = M;
        }
        const double prev[4] = {accumulate ? out[P_11 * half + g] : 0.0,
                                accumulate ? out[P_22 * half + g] : 0.0,
                                accumulate ? out[P_12R * half + g] : 0.0,
                                accumulate ? out[P_12I * half + g] : 0.0};
        out[P_PH * half + g] = 0.0;
        out[P_11 * half + g] = prev[0] + acc[i][0];
        out[P_22 * half + g] = prev[1] + acc[i][1];
        out[P_12R * half + g] = prev[2] + acc[i][2];
        out[P_12I * half + g] = prev[3] + acc[i][3];
    }
}

// The same second pass at M2 = 256 with two radix-16 passes per row: each
// wave transforms one frame's four rows (lane = 16 q + j: row q, butterfly
// j) -- the first pass from registers, its outputs multiplied by the second
// pass's twiddles W_256^{j r} (held per lane) and transposed through LDS,
// the second pass writing X[j + 16 r] -- so a frame costs 48 LDS accesses
// per lane instead of the radix-4 passes' 112 over four waves.  The four
// waves transform four frames, then after a barrier every thread forms its
// bins of all four (k_large_pass2_mac's bin code), and a second barrier
// frees the rows.
__device__ __forceinline__ void pair_bins(const float2 (*zb)[274], const float2 (*twp)[256],
                                          int tid, int M, int M1, int ra, int rb, int gs, int ov,
                                          int sbo, int jo, int grid, bool do_acf, bool do_xcf,
                                          bool ov_only, bool nyq, double dth, double phw,
                                          double (&acc)[3][4]) {
    constexpr int M2 = 256;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        int g, row, k2;
        if (i < 2) {
            row = i == 0 ? ra : rb;
            k2 = tid;
            g = row + M1 * k2;
        } else {
            if (!nyq) continue;
            row = 0;
            k2 = M2;
            g = M;
        }
        const int bg = row == ra ? 0 : 1;
        const int bm = (row == 0 ? row : M1 - row) == ra ? 0 : 1;
        const int km = row == 0 ? (M2 - k2) & (M2 - 1) : M2 - 1 - k2;
        const float2 Zg = zb[bg][pad(k2 & (M2 - 1))], Zm = zb[bm][pad(km)];
        const float2 Xg = untangle_ab(Zg, Zm, g, M, twp[bg][k2 & (M2 - 1)]);
        const bool in_ov = g >= gs && g < gs + ov;
        float2 Xo = make_float2(0.0f, 0.0f);
        if (in_ov) {
            const int h = g + sbo, kh = k2 + jo;
            const int khm = row == 0 ? (M2 - kh) & (M2 - 1) : M2 - 1 - kh;
            const float2 Oh = zb[2 + bg][pad(kh & (M2 - 1))];
            const float2 Om = zb[2 + bm][pad(khm & (M2 - 1))];
            Xo = untangle_ab(Oh, Om, h, M, twp[bg][kh & (M2 - 1)]);
        }
        if (do_acf) {
            const float pgrid = (!ov_only || in_ov) ? norm_sqr(Xg) : 0.0f;
            const float poth = in_ov ? norm_sqr(Xo) : 0.0f;
            acc[i][0] += (double)(grid == 0 ? pgrid : poth);
            acc[i][1] += (double)(grid == 0 ? poth : pgrid);
        }
        if (in_ov && do_xcf) {
            const float r = (float)wrap_turns(fma((double)(g - gs), dth, phw));
            const float2 phf = make_float2(__builtin_amdgcn_cosf(r), __builtin_amdgcn_sinf(r));
            const float2 X1 = (grid == 0) ? Xg : Xo;
            const float2 X2 = (grid == 0) ? Xo : Xg;
            const float2 v2 = c_mul(c_mul(X1, c_conj(X2)), phf);
            acc[i][2] += (double)v2.x;
            acc[i][3] += (double)v2.y;
        }
    }
}

__global__ __launch_bounds__(256) void k_large_pass2_r16(SectorArgs a, const float2 *y1,
                                                         const float2 *y2, const FrameParam *params,
                                                         int M1, const float2 *tw_m2, int64_t fpg,
                                                         int accumulate) {
    constexpr int M2 = 256, LP = M2 + M2 / 16 + 2;
    __shared__ float2 zr[4][4][LP];   // [frame of the iteration = wave][grid ra, grid rb, other ra, other rb]
    __shared__ float2 twp[2][M2];     // tw_post[row + M1 k2], rows ra / rb
    static_assert(LP == 274, "pair_bins row pitch");
    const int M = a.M;
    const int half = M + 1;
    const int tid = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const int q = lane >> 4, j = lane & 15;
    const int p = blockIdx.x;
    const int ra = p, rb = p == 0 ? M1 / 2 : M1 - p;
    for (int i = tid; i < 2 * M2; i += 256)
        twp[i / M2][i % M2] = a.tw_post[(i < M2 ? ra : rb) + M1 * (i % M2)];
    float2 tw[16];   // W_256^{j r}
#pragma unroll
    for (int r = 1; r < 16; ++r) tw[r] = tw_m2[(j * r) & (M2 - 1)];
    const int64_t g0 = (int64_t)blockIdx.y * fpg;
    int64_t g1 = g0 + fpg;
    if (g1 > a.nf) g1 = a.nf;
    const bool do_acf = (a.flags & FLAG_ACF) != 0;
    const bool do_xcf = (a.flags & FLAG_XCF) != 0;
    const bool ov_only = (a.flags & FLAG_ACF_OVERLAP) != 0;
    const int gs = (a.grid == 0) ? a.a1s : a.a2s;
    const int sbo = (a.grid == 0) ? (a.a2s - a.a1s) : (a.a1s - a.a2s);
    const int jo = sbo / M1;
    const float2 *src_row = ((q & 2) ? (a.grid == 0 ? y2 : y1) : (a.grid == 0 ? y1 : y2)) +
                            (int64_t)((q & 1) ? rb : ra) * M2 + j;
    const bool nyq = p == 0 && tid == 0;
    double acc[3][4];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[i][c] = 0.0;
    float2 cur[16];
    {
        const int64_t f = g0 + w;
#pragma unroll
        for (int r = 0; r < 16; ++r)
            cur[r] = f < g1 ? src_row[f * (int64_t)M + 16 * r] : make_float2(0.f, 0.f);
    }
    __syncthreads();   // twiddles staged
    float2 *z = zr[w][q];
    for (int64_t it = g0; it < g1; it += 4) {
        const int64_t f = it + w, fn = f + 4;
        float2 nxt[16];
#pragma unroll
        for (int r = 0; r < 16; ++r)
            nxt[r] = fn < g1 ? src_row[fn * (int64_t)M + 16 * r] : make_float2(0.f, 0.f);
        if (f < g1) {
            dft16(cur);                                  // pass 1: inputs j + 16 r
#pragma unroll
            for (int r = 1; r < 16; ++r) cur[r] = c_mul(cur[r], tw[r]);
#pragma unroll
            for (int r = 0; r < 16; ++r) z[pad(16 * j + r)] = cur[r];
            wave_sync_lds();
#pragma unroll
            for (int r = 0; r < 16; ++r) cur[r] = z[pad(j + 16 * r)];
            dft16(cur);                                  // pass 2
            wave_sync_lds();
#pragma unroll
            for (int r = 0; r < 16; ++r) z[pad(j + 16 * r)] = cur[r];   // X[j + 16 r]
        }
        __syncthreads();   // the iteration's rows are ready
        const int nfr = (int)((g1 - it) < 4 ? (g1 - it) : 4);
        for (int s2 = 0; s2 < nfr; ++s2)
            pair_bins(zr[s2], twp, tid, M, M1, ra, rb, gs, a.ov, sbo, jo, a.grid, do_acf, do_xcf,
                      ov_only, nyq, params[it + s2].dth, params[it + s2].phw, acc);
        __syncthreads();   // every thread is done with the rows
#pragma unroll
        for (int r = 0; r < 16; ++r) cur[r] = nxt[r];
    }
    // partial slot in pair-major order (as k_large_pass2_mac)
    double *out = a.part + (size_t)blockIdx.y * kNumProducts * half;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        int g;
        if (i < 2) {
            g = p * 2 * M2 + tid + 256 * i;
        } else {
            if (!nyq) continue;
            g = M;
        }
        const double prev[4] = {accumulate ? out[P_11 * half + g] : 0.0,
                                accumulate ? out[P_22 * half + g] : 0.0,
                                accumulate ? out[P_12R * half + g] : 0.0,
                                accumulate ? out[P_12I * half + g] : 0.0};
        out[P_PH * half + g] = 0.0;
        out[P_11 * half + g] = prev[0] + acc[i][0];
        out[P_22 * half + g] = prev[1] + acc[i][1];
        out[P_12R * half + g] = prev[2] + acc[i][2];
        out[P_12I * half + g] = prev[3] + acc[i][3];
    }
}

// One bin of a row pair for the second pass at any M2: bin g = row + M1 k2
// of row ra (rb_sel false) or rb, or bin M (nyq: row 0, k2 = M2); the
// untangle of both stations and the ACF/XCF products (pair_bins' arithmetic).
template <int M2, int LP>
__device__ __forceinline__ void row_bin(const float2 (*zb)[LP], const float2 (*twp)[M2], bool rb_sel,
                                        int k2, bool nyq, int M, int M1, int ra, int rb, int gs, int ov,
                                        int sbo, int jo, int grid, bool do_acf, bool do_xcf,
                                        bool ov_only, double dth, double phw, double (&acc)[4]) {
    const int row = nyq ? 0 : (rb_sel ? rb : ra);
    if (nyq) k2 = M2;
    const int g = nyq ? M : row + M1 * k2;
    const int bg = row == ra ? 0 : 1;
    const int bm = (row == 0 ? row : M1 - row) == ra ? 0 : 1;
    const int km = row == 0 ? (M2 - k2) & (M2 - 1) : M2 - 1 - k2;
    const float2 Zg = zb[bg][pad(k2 & (M2 - 1))], Zm = zb[bm][pad(km)];
    const float2 Xg = untangle_ab(Zg, Zm, g, M, twp[bg][k2 & (M2 - 1)]);
    const bool in_ov = g >= gs && g < gs + ov;
    float2 Xo = make_float2(0.0f, 0.0f);
    if (in_ov) {
        const int h = g + sbo, kh = k2 + jo;
        const int khm = row == 0 ? (M2 - kh) & (M2 - 1) : M2 - 1 - kh;
        const float2 Oh = zb[2 + bg][pad(kh & (M2 - 1))];
        const float2 Om = zb[2 + bm][pad(khm & (M2 - 1))];
        Xo = untangle_ab(Oh, Om, h, M, twp[bg][kh & (M2 - 1)]);
    }
    if (do_acf) {
        const float pgrid = (!ov_only || in_ov) ? norm_sqr(Xg) : 0.0f;
        const float poth = in_ov ? norm_sqr(Xo) : 0.0f;
        acc[0] += (double)(grid == 0 ? pgrid : poth);
        acc[1] += (double)(grid == 0 ? poth : pgrid);
    }
    if (in_ov && do_xcf) {
        const float r = (float)wrap_turns(fma((double)(g - gs), dth, phw));
        const float2 phf = make_float2(__builtin_amdgcn_cosf(r), __builtin_amdgcn_sinf(r));
        const float2 X1 = (grid == 0) ? Xg : Xo;
        const float2 X2 = (grid == 0) ? Xo : Xg;
        const float2 v2 = c_mul(c_mul(X1, c_conj(X2)), phf);
        acc[2] += (double)v2.x;
        acc[3] += (double)v2.y;
    }
}

// The second pass at M2 = 64 (N = 16384, M1 = 128) with two radix-8 passes
// per row: lane 8 q + j of wave w takes row slot q (frame 2 w + q / 4 of the
// iteration, row q % 4 of [grid ra, grid rb, other ra, other rb]) and
// butterfly j -- the first pass from registers (inputs j + 8 r), its outputs
// multiplied by W_64^{j r} and transposed through LDS, the second pass
// writing X[j + 8 r] -- so the four waves transform eight frames per
// iteration.  After a barrier thread t forms bin t % 128 of the pair (row ra
// for t % 128 < 64, else rb; k2 = t % 64) for frames t / 128, + 2, + 4, + 6
// of the eight, and a second barrier frees the rows.  Threads t and t + 128
// hold the same bin; their sums are added in that order at the end (bin M,
// pair 0: threads 0 and 128).
__global__ __launch_bounds__(256) void k_large_pass2_r8(SectorArgs a, const float2 *y1,
                                                        const float2 *y2, const FrameParam *params,
                                                        int M1, const float2 *tw_m2, int64_t fpg,
                                                        int accumulate) {
    constexpr int M2 = 64, LP = M2 + M2 / 16 + 2;
    __shared__ float2 zr[8][4][LP];   // [frame of the iteration][grid ra, grid rb, other ra, other rb]
    __shared__ float2 twp[2][M2];     // tw_post[row + M1 k2], rows ra / rb
    __shared__ double hsum[128][4];   // threads 128..255's sums
    __shared__ double msum[4];        // thread 128's bin-M sums (pair 0)
    const int M = a.M;
    const int half = M + 1;
    const int tid = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const int q = lane >> 3, j = lane & 7;
    const int fw = q >> 2, rk = q & 3;
    const int p = blockIdx.x;
    const int ra = p, rb = p == 0 ? M1 / 2 : M1 - p;
    for (int i = tid; i < 2 * M2; i += 256)
        twp[i / M2][i % M2] = a.tw_post[(i < M2 ? ra : rb) + M1 * (i % M2)];
    float2 tw[8];
#pragma unroll
    for (int r = 1; r < 8; ++r) tw[r] = tw_m2[(j * r) & (M2 - 1)];
    const int64_t g0 = (int64_t)blockIdx.y * fpg;
    int64_t g1 = g0 + fpg;
    if (g1 > a.nf) g1 = a.nf;
    const bool do_acf = (a.flags & FLAG_ACF) != 0;
    const bool do_xcf = (a.flags & FLAG_XCF) != 0;
    const bool ov_only = (a.flags & FLAG_ACF_OVERLAP) != 0;
    const int gs = (a.grid == 0) ? a.a1s : a.a2s;
    const int sbo = (a.grid == 0) ? (a.a2s - a.a1s) : (a.a1s - a.a2s);
    const int jo = sbo / M1;
    const float2 *src_row = ((rk & 2) ? (a.grid == 0 ? y2 : y1) : (a.grid == 0 ? y1 : y2)) +
                            (int64_t)((rk & 1) ? rb : ra) * M2 + j;
    const int bsel = tid & 127, fsel = tid >> 7;
    const bool rb_sel = bsel >= M2;
    const int k2 = bsel & (M2 - 1);
    const bool nyq = p == 0 && bsel == 0;   // threads 0 and 128 also take bin M
    double acc[4] = {0.0, 0.0, 0.0, 0.0}, nacc[4] = {0.0, 0.0, 0.0, 0.0};
    const int slot = 2 * w + fw;   // this lane's frame of the iteration
    float2 cur[8];
    {
        const int64_t f = g0 + slot;
#pragma unroll
        for (int r = 0; r < 8; ++r)
            cur[r] = f < g1 ? src_row[f * (int64_t)M + 8 * r] : make_float2(0.f, 0.f);
    }
    __syncthreads();   // twiddles staged
    float2 *z = zr[slot][rk];
    for (int64_t it = g0; it < g1; it += 8) {
        const int64_t f = it + slot, fn = f + 8;
        float2 nxt[8];
#pragma unroll
        for (int r = 0; r < 8; ++r)
            nxt[r] = fn < g1 ? src_row[fn * (int64_t)M + 8 * r] : make_float2(0.f, 0.f);
        if (f < g1) {
            dftR<8>(cur);                                // pass 1: inputs j + 8 r
#pragma unroll
            for (int r = 1; r < 8; ++r) cur[r] = c_mul(cur[r], tw[r]);
#pragma unroll
            for (int r = 0; r < 8; ++r) z[pad(8 * j + r)] = cur[r];
            wave_sync_lds();
#pragma unroll
            for (int r = 0; r < 8; ++r) cur[r] = z[pad(j + 8 * r)];
            dftR<8>(cur);                                // pass 2
            wave_sync_lds();
#pragma unroll
            for (int r = 0; r < 8; ++r) z[pad(j + 8 * r)] = cur[r];   // X[j + 8 r]
        }
        __syncthreads();   // the iteration's rows are ready
        const int nfr = (int)((g1 - it) < 8 ? (g1 - it) : 8);
        for (int s2 = fsel; s2 < nfr; s2 += 2) {
            const double dth = params[it + s2].dth, phw = params[it + s2].phw;
            row_bin<M2, LP>(zr[s2], twp, rb_sel, k2, false, M, M1, ra, rb, gs, a.ov, sbo, jo, a.grid,
                            do_acf, do_xcf, ov_only, dth, phw, acc);
            if (nyq)
                row_bin<M2, LP>(zr[s2], twp, false, 0, true, M, M1, ra, rb, gs, a.ov, sbo, jo, a.grid,
                                do_acf, do_xcf, ov_only, dth, phw, nacc);
        }
        __syncthreads();   // every thread is done with the rows
#pragma unroll
        for (int r = 0; r < 8; ++r) cur[r] = nxt[r];
    }
    if (tid >= 128) {
#pragma unroll
        for (int c = 0; c < 4; ++c) hsum[bsel][c] = acc[c];
        if (nyq)
#pragma unroll
            for (int c = 0; c < 4; ++c) msum[c] = nacc[c];
    }
    __syncthreads();
    if (tid >= 128) return;
    // partial slot in pair-major order (as k_large_pass2_mac)
    double *out = a.part + (size_t)blockIdx.y * kNumProducts * half;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        int g;
        double v[4];
        if (i == 0) {
            g = p * 2 * M2 + bsel;
#pragma unroll
            for (int c = 0; c < 4; ++c) v[c] = acc[c] + hsum[bsel][c];
        } else {
            if (!nyq) continue;
            g = M;
#pragma unroll
            for (int c = 0; c < 4; ++c) v[c] = nacc[c] + msum[c];
        }
        const double prev[4] = {accumulate ? out[P_11 * half + g] : 0.0,
                                accumulate ? out[P_22 * half + g] : 0.0,
                                accumulate ? out[P_12R * half + g] : 0.0,
                                accumulate ? out[P_12I * half + g] : 0.0};
        out[P_PH * half + g] = 0.0;
        out[P_11 * half + g] = prev[0] + v[0];
        out[P_22 * half + g] = prev[1] + v[1];
        out[P_12R * half + g] = prev[2] + v[2];
        out[P_12I * half + g] = prev[3] + v[3];
    }
}

// One bin of a row pair for the phased sum (k_phased_large's per-bin
// arithmetic, main.rs:2264-2302): both stations' untangled bins, each
// rotated by exp(2 pi i b th) from the FP64 phase and by its fringe phasor,
// cb = w1 c1 + w2 c2, |cb|^2 into the FP64 sum and, with cbl, cb (DC and
// Nyquist imaginary parts zeroed) into the frame's pair rows in LDS
// (cbl[row ra / rb][k2], bin M in *cbn).
template <int M2, int LP>
__device__ __forceinline__ void row_bin_ph(const float2 (*zb)[LP], const float2 (*twp)[M2], bool rb_sel,
                                           int k2, bool nyq, int M, int M1, int ra, int rb, int gs,
                                           int ov, int sbo, int jo, int grid, const PhasedParam &P,
                                           float w1, float w2, float2 (*cbl)[M2], float2 *cbn,
                                           double &acc) {
    const int row = nyq ? 0 : (rb_sel ? rb : ra);
    if (nyq) k2 = M2;
    const int g = nyq ? M : row + M1 * k2;
    const int bg = row == ra ? 0 : 1;
    const int bm = (row == 0 ? row : M1 - row) == ra ? 0 : 1;
    const int km = row == 0 ? (M2 - k2) & (M2 - 1) : M2 - 1 - k2;
    const float2 Zg = zb[bg][pad(k2 & (M2 - 1))], Zm = zb[bm][pad(km)];
    const float2 Xg = untangle_ab(Zg, Zm, g, M, twp[bg][k2 & (M2 - 1)]);
    const bool in_ov = g >= gs && g < gs + ov;
    const double thg = grid == 0 ? P.th1 : P.th2, tho = grid == 0 ? P.th2 : P.th1;
    const float2 frg = grid == 0 ? P.fr1 : P.fr2, fro = grid == 0 ? P.fr2 : P.fr1;
    float2 sg = c_mul(Xg, phasor_turns((double)g * thg));
    if (g == M) sg.y = 0.0f;                              // Nyquist (utils.rs:118-123)
    const float2 cg = c_mul(sg, frg);
    float2 co = make_float2(0.0f, 0.0f);
    if (in_ov) {
        const int h = g + sbo, kh = k2 + jo;
        const int khm = row == 0 ? (M2 - kh) & (M2 - 1) : M2 - 1 - kh;
        const float2 Oh = zb[2 + bg][pad(kh & (M2 - 1))];
        const float2 Om = zb[2 + bm][pad(khm & (M2 - 1))];
        float2 so = c_mul(untangle_ab(Oh, Om, h, M, twp[bg][kh & (M2 - 1)]), phasor_turns((double)h * tho));
        if (h == M) so.y = 0.0f;
        co = c_mul(so, fro);
    }
    const float2 c1 = grid == 0 ? cg : co, c2 = grid == 0 ? co : cg;
    float2 cb;
    {
#pragma clang fp contract(off)
        const float2 a1 = c_scale(c1, w1), a2 = c_scale(c2, w2);
        cb = make_float2(a1.x + a2.x, a1.y + a2.y);
    }
    acc += (double)norm_sqr(cb);
    if (cbl) {
        if (g == 0 || g == M) cb.y = 0.0f;                // main.rs:2283-2284
        if (nyq) *cbn = cb;
        else cbl[bg][k2] = cb;
    }
}

// The phased sum on the second pass at M2 = 64 (N = 16384): k_large_pass2_r8
// with the phased bin in place of the ACF/XCF products.  RAW (the .raw
// output): the first half of the inverse real FFT as well.  The pair's cb
// rows of the iteration's frames are kept in LDS; after the barrier that
// ends the bins, lanes 0..31 of wave w take (frame, row) tasks 4 w ..
// 4 w + 3 (frame task / 2, row ra or rb), 8 lanes each: y[k] = conj(Z'[k])
// of the half-length complex inverse (irfft_pre_elem's arithmetic: bins k
// and M - k lie in the pair), the row's DFT-64 over k2 (two radix-8 passes,
// the wave's own row buffers as scratch), times W_M^{n2 k1}, stored as row
// k1 of the frame's 128 x 64 block in a.cb_out (M entries per frame).
// k_irfft_cols_quantise finishes the transform over k1 and quantises.
template <bool RAW>
__global__ __launch_bounds__(256, RAW ? 4 : 5) void k_large_pass2_r8_ph(SectorArgs a, const float2 *y1,
                                                           const float2 *y2, const PhasedParam *pp,
                                                           int M1, const float2 *tw_m2,
                                                           const float2 *tw_m, int64_t fpg,
                                                           int accumulate) {
    constexpr int M2 = 64, LP = M2 + M2 / 16 + 2;
    __shared__ float2 zr[8][4][LP];   // [frame of the iteration][grid ra, grid rb, other ra, other rb]
    __shared__ float2 twp[2][M2];     // tw_post[row + M1 k2], rows ra / rb
    __shared__ double hsum[128];      // threads 128..255's sums
    __shared__ double msum;           // thread 128's bin-M sum (pair 0)
    __shared__ float2 cbl[RAW ? 8 : 1][2][M2];   // cb of the pair's rows, per frame
    __shared__ float2 cbn[8];                    // cb of bin M (pair 0)
    __shared__ float2 twr[RAW ? 2 : 1][M2];      // W_M^{n2 row}, rows ra / rb
    const int M = a.M;
    const int half = M + 1;
    const int tid = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const int q = lane >> 3, j = lane & 7;
    const int fw = q >> 2, rk = q & 3;
    const int p = blockIdx.x;
    const int ra = p, rb = p == 0 ? M1 / 2 : M1 - p;
    for (int i = tid; i < 2 * M2; i += 256) {
        twp[i / M2][i % M2] = a.tw_post[(i < M2 ? ra : rb) + M1 * (i % M2)];
        if constexpr (RAW) twr[i / M2][i % M2] = tw_m[((i % M2) * (i < M2 ? ra : rb)) & (M - 1)];
    }
    float2 tw[8];
#pragma unroll
    for (int r = 1; r < 8; ++r) tw[r] = tw_m2[(j * r) & (M2 - 1)];
    const int64_t g0 = (int64_t)blockIdx.y * fpg;
    int64_t g1 = g0 + fpg;
    if (g1 > a.nf) g1 = a.nf;
    const float w1 = (float)a.w1, w2 = (float)a.w2;
    const int gs = (a.grid == 0) ? a.a1s : a.a2s;
    const int sbo = (a.grid == 0) ? (a.a2s - a.a1s) : (a.a1s - a.a2s);
    const int jo = sbo / M1;
    const float2 *src_row = ((rk & 2) ? (a.grid == 0 ? y2 : y1) : (a.grid == 0 ? y1 : y2)) +
                            (int64_t)((rk & 1) ? rb : ra) * M2 + j;
    const int bsel = tid & 127, fsel = tid >> 7;
    const bool rb_sel = bsel >= M2;
    const int k2 = bsel & (M2 - 1);
    const bool nyq = p == 0 && bsel == 0;   // threads 0 and 128 also take bin M
    double acc = 0.0, nacc = 0.0;
    const int slot = 2 * w + fw;   // this lane's frame of the iteration
    float2 cur[8];
    {
        const int64_t f = g0 + slot;
#pragma unroll
        for (int r = 0; r < 8; ++r)
            cur[r] = f < g1 ? src_row[f * (int64_t)M + 8 * r] : make_float2(0.f, 0.f);
    }
    __syncthreads();   // twiddles staged
    float2 *z = zr[slot][rk];
    for (int64_t it = g0; it < g1; it += 8) {
        const int64_t f = it + slot, fn = f + 8;
        float2 nxt[8];
#pragma unroll
        for (int r = 0; r < 8; ++r)
            nxt[r] = fn < g1 ? src_row[fn * (int64_t)M + 8 * r] : make_float2(0.f, 0.f);
        if (f < g1) {
            dftR<8>(cur);                                // pass 1: inputs j + 8 r
#pragma unroll
            for (int r = 1; r < 8; ++r) cur[r] = c_mul(cur[r], tw[r]);
#pragma unroll
            for (int r = 0; r < 8; ++r) z[pad(8 * j + r)] = cur[r];
            wave_sync_lds();
#pragma unroll
            for (int r = 0; r < 8; ++r) cur[r] = z[pad(j + 8 * r)];
            dftR<8>(cur);                                // pass 2
            wave_sync_lds();
#pragma unroll
            for (int r = 0; r < 8; ++r) z[pad(j + 8 * r)] = cur[r];   // X[j + 8 r]
        }
        __syncthreads();   // the iteration's rows are ready
        const int nfr = (int)((g1 - it) < 8 ? (g1 - it) : 8);
        for (int s2 = fsel; s2 < nfr; s2 += 2) {
            const PhasedParam P = pp[it + s2];
            float2(*cbf)[M2] = RAW ? cbl[RAW ? s2 : 0] : nullptr;
            row_bin_ph<M2, LP>(zr[s2], twp, rb_sel, k2, false, M, M1, ra, rb, gs, a.ov, sbo, jo,
                               a.grid, P, w1, w2, cbf, cbn + s2, acc);
            if (nyq)
                row_bin_ph<M2, LP>(zr[s2], twp, false, 0, true, M, M1, ra, rb, gs, a.ov, sbo, jo,
                                   a.grid, P, w1, w2, cbf, cbn + s2, nacc);
        }
        __syncthreads();   // every thread is done with the rows; cb complete
        if constexpr (RAW) {
            // the wave's row buffers zr[2 w][*] are scratch here: the next
            // frames' rows are written to them by this wave only
            const int task = 4 * w + (lane >> 3);
            const int s3 = task >> 1, b = task & 1;
            if (lane < 32 && s3 < nfr) {
                const int row = b ? rb : ra;
                const int bm = (row == 0 ? row : M1 - row) == ra ? 0 : 1;
                float2 v[8];
#pragma unroll
                for (int r = 0; r < 8; ++r) {
                    const int kk = j + 8 * r;                  // k2 of bin k = row + M1 kk
                    float2 zq;
                    if (row == 0 && kk == 0) {
                        const float2 x0 = cbl[s3][0][0], xm = cbn[s3];
                        zq = make_float2(x0.x + xm.x, x0.x - xm.x);
                    } else {
                        const int km = row == 0 ? (M2 - kk) & (M2 - 1) : M2 - 1 - kk;
                        const float2 A = cbl[s3][b][kk], Bc = cbl[s3][bm][km];
                        const float er = A.x + Bc.x, ei = A.y - Bc.y;
                        const float dr = A.x - Bc.x, di = A.y + Bc.y;
                        const float2 wp = twp[b][kk];
                        const float orr = dr * wp.x - di * wp.y;
                        const float oi = dr * wp.y + di * wp.x;
                        zq = make_float2(er - oi, ei + orr);
                    }
                    v[r] = c_conj(zq);
                }
                float2 *zs = zr[2 * w][lane >> 3];
                dftR<8>(v);                                  // pass 1: inputs j + 8 r
#pragma unroll
                for (int r = 1; r < 8; ++r) v[r] = c_mul(v[r], tw[r]);
#pragma unroll
                for (int r = 0; r < 8; ++r) zs[pad(8 * j + r)] = v[r];
                wave_sync_lds();
#pragma unroll
                for (int r = 0; r < 8; ++r) v[r] = zs[pad(j + 8 * r)];
                dftR<8>(v);                                  // pass 2: outputs n2 = j + 8 r
                float2 *bo = a.cb_out + (it + s3) * (int64_t)M + row * M2;
#pragma unroll
                for (int r = 0; r < 8; ++r) bo[j + 8 * r] = c_mul(v[r], twr[b][j + 8 * r]);
                wave_sync_lds();
            }
        }
#pragma unroll
        for (int r = 0; r < 8; ++r) cur[r] = nxt[r];
    }
    if (tid >= 128) {
        hsum[bsel] = acc;
        if (nyq) msum = nacc;
    }
    __syncthreads();
    if (tid >= 128) return;
    // partial slot in pair-major order (as k_large_pass2_mac); the other
    // products stay zero
    double *out = a.part + (size_t)blockIdx.y * kNumProducts * half;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        int g;
        double v;
        if (i == 0) {
            g = p * 2 * M2 + bsel;
            v = acc + hsum[bsel];
        } else {
            if (!nyq) continue;
            g = M;
            v = nacc + msum;
        }
        out[P_PH * half + g] = accumulate ? out[P_PH * half + g] + v : v;
        out[P_11 * half + g] = 0.0;
        out[P_22 * half + g] = 0.0;
        out[P_12R * half + g] = 0.0;
        out[P_12I * half + g] = 0.0;
    }
}

hipError_t launch_large_pass2_phased(const SectorArgs &a, const float2 *y1, const float2 *y2,
                                     const void *pp, int M1, int M2, const float2 *tw_m2,
                                     const float2 *tw_m, bool raw, int n_groups, int accumulate,
                                     hipStream_t s) {
    const int sbo = (a.grid == 0) ? (a.a2s - a.a1s) : (a.a1s - a.a2s);
    if (M1 * M2 != a.M || sbo % M1 != 0 || M2 != 64 || M1 != 128 || (raw && !a.cb_out))
        return hipErrorInvalidValue;
    const int64_t fpg = (a.nf + n_groups - 1) / n_groups;
    const dim3 grid((unsigned)(M1 / 2), (unsigned)n_groups);
    if (raw)
        hipLaunchKernelGGL(k_large_pass2_r8_ph<true>, grid, dim3(256), 0, s, a, y1, y2,
                           (const PhasedParam *)pp, M1, tw_m2, tw_m, fpg, accumulate);
    else
        hipLaunchKernelGGL(k_large_pass2_r8_ph<false>, grid, dim3(256), 0, s, a, y1, y2,
                           (const PhasedParam *)pp, M1, tw_m2, tw_m, fpg, accumulate);
    return hipGetLastError();
}

hipError_t launch_large_pass2_mac(const SectorArgs &a, const float2 *y1, const float2 *y2,
                                  const FrameParam *params, int M1, int M2, const float2 *tw_m2,
                                  int n_groups, int accumulate, hipStream_t s) {
    const int sbo = (a.grid == 0) ? (a.a2s - a.a1s) : (a.a1s - a.a2s);
    if (M1 * M2 != a.M || sbo % M1 != 0 || (M2 != 64 && M2 != 128 && M2 != 256)) return hipErrorInvalidValue;
    const int64_t fpg = (a.nf + n_groups - 1) / n_groups;
    const dim3 grid((unsigned)(M1 / 2), (unsigned)n_groups);
    if (M2 == 64)
        hipLaunchKernelGGL(k_large_pass2_r8, grid, dim3(256), 0, s, a, y1, y2, params, M1, tw_m2, fpg,
                           accumulate);
    else if (M2 == 128)
        hipLaunchKernelGGL(k_large_pass2_mac<128>, grid, dim3(256), 0, s, a, y1, y2, params, M1, tw_m2,
                           fpg, accumulate);
    else
        hipLaunchKernelGGL(k_large_pass2_r16, grid, dim3(256), 0, s, a, y1, y2, params, M1, tw_m2,
                           fpg, accumulate);
    return hipGetLastError();
}

hipError_t launch_mac_large(const SectorArgs &a, int precision, const float2 *z1, const float2 *z2,
                            const FrameParam *params, int n_groups, int accumulate,
                            hipStream_t s) {
    const int64_t fpg = (a.nf + n_groups - 1) / n_groups;
    const dim3 grid((unsigned)((a.M + 1 + 255) / 256), (unsigned)n_groups);
    if (precision == 0)
        hipLaunchKernelGGL(k_mac_large<0>, grid, dim3(256), 0, s, a, z1, z2, params, fpg, accumulate);
    else
        hipLaunchKernelGGL(k_mac_large<1>, grid, dim3(256), 0, s, a, z1, z2, params, fpg, accumulate);
    return hipGetLastError();
}

}  // namespace fxc
