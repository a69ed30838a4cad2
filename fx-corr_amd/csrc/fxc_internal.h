// Internal declarations shared by the C-ABI layer (fxc_api.cpp) and the
// device code (fxc_device.hip).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace fxc {

constexpr int kMaxBits = 16;       // widest sample the GPU decoder accepts
constexpr int kFrameBatch = 8;     // frames whose delay entries one batch computes
constexpr int kMaxPow = 16;        // step^(2^b), b < kMaxPow (exponents below 65536)
constexpr int kNumProducts = 5;    // [ph | 11 | 22 | 12re | 12im]
enum { P_PH = 0, P_11 = 1, P_22 = 2, P_12R = 3, P_12I = 4 };

enum { SHUF_IDENTITY = 0, SHUF_PAIRSWAP = 1, SHUF_GENERIC = 2 };
enum { FLAG_ACF = 1, FLAG_XCF = 2, FLAG_PHASED = 4, FLAG_RAW_OUT = 8, FLAG_ACF_OVERLAP = 16 };

// One station's decode description, device side (utils.rs:147-155).
struct StationDev {
    int32_t bits;
    int32_t lsb;
    int32_t shuffle_kind;
    int32_t bpf;          // bytes per frame
    int32_t nwords;       // 32-bit words staged per frame (ceil(bpf/4))
    int32_t aligned;      // bpf % 4 == 0: frames start on a word boundary
    int32_t decode_ok;    // plan-level: every frame decodes to N samples
    uint32_t code_mask;
    int32_t shifts[32];   // out bit i <- in bit shifts[i]
    float lev4[16];       // levels for bits <= 4 (kernel-argument resident)
    const float *levels;  // device table, 1<<bits entries
};

// DelayEvalConfig (main.rs:161-182), device side.
struct DelayDev {
    double frame_dt, fs, time_offset_s;
    const double *table;
    int64_t table_len;
    double coarse, user_samples, extra_rate;
    double c1d, c1r, c2d, c2r;
    double d_seek;
    double nd1, tr1, ta1, nd2, tr2, ta2;
    double lo1, lo2;
};

// FrameDelayEntry (main.rs:151-159), device side.
struct FrameDelayDev {
    int64_t int1, int2;
    double frac1, frac2;
    float fr1[2], fr2[2];
};

struct SectorArgs {
    const uint8_t *raw1;
    const uint8_t *raw2;
    int64_t first_frame;     // global index of frame 0 of this launch
    int64_t nf;              // frames in this launch
    int64_t frames_per_wg;
    StationDev st[2];
    DelayDev dly;
    const float2 *tw_fft;    // W_M^k, k < M
    const float2 *tw_post;   // (cos, sin)(2*pi*k/N), k <= M
    int32_t N, M;
    int32_t a1s, a1e, a2s, ov;
    int32_t grid;            // 0 = Ant1, 1 = Ant2
    int32_t flags;
    double df;
    double w1, w2;
    int32_t bit_out;
    int32_t output_lsb;
    int32_t out_shifts[32];  // station-1 shuffle used on output words
    int32_t out_shuffle_ident;  // 1: out_shifts is the identity
    const float *out_levels; // station-1 levels (f32 as the reference compares)
    int32_t out_nlevels;
    int32_t out_sorted;      // 1: out_levels non-decreasing (nearest level by bisection)
    int32_t out_qmode;       // quantiser search: 0 scan, 1 bisection, 2 spacing guess
    float out_l0, out_inv_step;  // mode 2: first level, 1 / mean level spacing
    int32_t bpf_out;
    uint8_t *raw_out;        // phased .raw frames (nf * bpf_out) or null
    double *part;            // [n_wg][kNumProducts][half]
    const void *mf_tab;      // chosen-form MFMA tables, layout MFT_* (N = 1024) or MFT4_* (N = 2048)
    float2 *cb_out;          // phased: combined half spectra [frame][half] for the .raw output
    int32_t variant;         // FFT pass sequence of the plain kernel (autotuned; 0 default)
    // Plain form: the reference's sequential f32 phase chains of this
    // launch's frames, row (frame * chains_per_frame + c) of chain_ld float2
    // (k_phase_chain); null where the kernels take FP64 powers of the step.
    const float2 *chain;
    int32_t chain_ld;
    int32_t accumulate;      // fused plain kernel: add to the partial slots (sector in slabs)
    int32_t *err;            // plan's device error word (bit 0: a bounded hand-over wait ran out)
    // N = 1024 producer/consumer kernel: per-consumer FP64 totals,
    // [n_wg][kMfConsumers][16][64] double2 (kMfConsumerAccBytes per workgroup)
    double *cons_acc;
};
constexpr int kMfConsumers = 4;
constexpr size_t kMfConsumerAccBytes = (size_t)kMfConsumers * 32 * 64 * sizeof(double);

// Phase chains replayed by k_phase_chain (fxc_chain.hip).
enum { CHAIN_XCF = 0, CHAIN_PHASED = 1 };
// The plain XCF replays the chain above this many overlap bins.  Up to it,
// FP64 powers of the same f32 step stay within 1.4e-5 of the chain over a
// frame's bins (2.8e-6 rms; 2.6e-5 and 6.6e-6 at 1024 bins), which keeps the
// fused N <= 1024 kernel free of a per-frame table.
constexpr int kChainMinOverlap = 512;
__host__ __device__ constexpr int chain_ld_for(int len) { return (len + 1) & ~1; }

// Chosen-form MFMA tables (fxc_mfma.hip), byte offsets into one device
// buffer.  [idx][lane] order so a wave's loads are contiguous.
constexpr int MFT_B1 = 0;                      // 16 x 64 uint4: DFT-32 FP16 hi/lo operands
constexpr int MFT_A2 = MFT_B1 + 16 * 64 * 16;  // 16 x 64 f32: DFT-16 A operands
constexpr int MFT_TW2 = MFT_A2 + 16 * 64 * 4;  // 16 x 64 f32: W512^{m2 k1}
constexpr int MFT_TWU = MFT_TW2 + 16 * 64 * 4; // 8 x 64 f32: W1024^p of the untangle pairs
constexpr int MFT_DLUT_HALVES = 2048;          // per station: byte -> FP16 samples
constexpr int MFT_DLUT = MFT_TWU + 8 * 64 * 4;
constexpr int MFT_SLUT = MFT_DLUT + 2 * MFT_DLUT_HALVES * 2;  // 2 x 4 x 256 u32
constexpr int MFT_BYTES = MFT_SLUT + 2 * 4 * 256 * 4;
// The same tables for N = 2048 (k_sector_mf4): the DFT-32 operands are
// those of N = 1024; the twiddles cover 32 rows m2 and 8 untangle pairs.
constexpr int MFT4_B1 = 0;                       // 16 x 64 uint4
constexpr int MFT4_TW2 = MFT4_B1 + 16 * 64 * 16; // 32 x 64 f32: W1024^{m2 k1}
constexpr int MFT4_TWU = MFT4_TW2 + 32 * 64 * 4; // 16 x 64 f32: W2048^p
constexpr int MFT4_DLUT = MFT4_TWU + 16 * 64 * 4;
constexpr int MFT4_SLUT = MFT4_DLUT + 2 * MFT_DLUT_HALVES * 2;
constexpr int MFT4_BYTES = MFT4_SLUT + 2 * 4 * 256 * 4;

// Launchers implemented in fxc_device.hip.  All enqueue on `stream`.
hipError_t launch_sector(const SectorArgs &a, int precision, int n_wg,
                         hipStream_t stream);
hipError_t launch_reduce(const double *part, int n_wg, int half, double *out,
                         double *scratch, int n_chunks, hipStream_t stream);
int sector_threads_per_frame(int N);
bool sector_supported(int N);
bool sector_has_variant(int N);

// out: nframes x N floats, or with fp16 nframes x N/2 FP16 pairs (dwords)
hipError_t launch_decode_frames(const SectorArgs &a, int station,
                                const int64_t *shifts, int shift_stride, void *out, bool fp16,
                                hipStream_t stream);
hipError_t launch_rfft_frames(const SectorArgs &a, const float *in, int64_t nframes,
                              float2 *out, hipStream_t stream);
hipError_t launch_irfft_frames(const SectorArgs &a, const float2 *in,
                               int64_t nframes, float *out, bool negate_odd,
                               hipStream_t stream);
hipError_t launch_delay_eval(const DelayDev &d, int64_t first, int64_t n,
                             FrameDelayDev *out, hipStream_t stream);
// Inverse FFT and quantiser in one kernel (phased .raw output, N <= 16384).
hipError_t launch_irfft_quantise(const SectorArgs &a, const float2 *in, int64_t nframes,
                                 uint8_t *out, bool negate_odd, hipStream_t s);
// Quantiser over the HBM four-step inverse's complex output (N >= 32768).
hipError_t launch_quantise_cplx(const SectorArgs &a, const float2 *z, int64_t nframes,
                                bool negate_odd, uint8_t *out, hipStream_t s);
hipError_t launch_quantise_frames(const SectorArgs &a, const float *in,
                                  int64_t nframes, uint8_t *out,
                                  hipStream_t stream);

hipError_t launch_sector_mfma(const SectorArgs &a, int precision, int n_wg, hipStream_t stream);
int mfma_wg_threads();
// Chosen phased sum at N = 256 / 512 / 1024 on the producer/consumer MFMA kernel (a1s == a2s)
hipError_t launch_phased_mf3(const SectorArgs &a, int accumulate, int n_wg, hipStream_t stream);

// fxc_phased.hip
// precision: 0 plain (reference-faithful f32 chains), 1 chosen (FP64 phases, no chains)
hipError_t launch_phased(const SectorArgs &a, int precision, int accumulate, int n_wg,
                         hipStream_t stream);
hipError_t launch_fill_zero(void *p, size_t bytes, hipStream_t stream);
hipError_t launch_add_f64(double *dst, const double *src, size_t n, hipStream_t stream);
hipError_t launch_scale_add_xcf(double *ac, const double *re, const double *im, int half, double s,
                                hipStream_t stream);

// fxc_chain.hip: one thread per chain; frames [0, a.nf) of the launch,
// rows of a.chain_ld float2 (one XCF chain of a.ov bins, or two phased
// chains of M+1 bins, per frame).
hipError_t launch_phase_chain(const SectorArgs &a, int kind, float2 *out, hipStream_t stream);

// fxc_large.hip: N = 32768, 65536 through HBM scratch (four-step M = M1 x M2)
struct FrameParam;
constexpr int kLargeGroups = 16;   // frame groups of the bin-owned product kernel
hipError_t launch_frame_params(const SectorArgs &a, int precision, FrameParam *params,
                               int64_t *shifts, hipStream_t stream);
// mf_b1: the FP16 hi/lo DFT-32 operand table for a first pass on MFMA
// (chosen form, inputs exact in FP16), or null for the FP32 pass.
// x_fp16: x holds FP16 sample pairs (N/2 dwords per frame; only with mf_b1)
hipError_t launch_fft_large(const void *x, bool x_fp16, float2 *ybuf, float2 *z, int64_t nframes,
                            int M1, int M2, const float2 *tw1, const float2 *tw2,
                            const float2 *twm, const void *mf_b1, hipStream_t stream);
hipError_t launch_fft_pass1_mfma(const void *in, bool in_fp16, float2 *out, int64_t nbatch, int M1,
                                 int M2, const void *b1, const float2 *tw_m, hipStream_t stream);
// inverse real FFT of nframes half spectra X through the HBM four-step:
// zc (M complex per frame), ybuf, z scratch; out nframes*N floats, or null
// to leave the unscaled transform (conj-FFT input, M per frame) in z.
hipError_t launch_irfft_large(const float2 *X, int64_t nframes, int M1, int M2, const float2 *tw1,
                              const float2 *tw2, const float2 *twm, const float2 *tw_post,
                              float2 *zc, float2 *ybuf, float2 *z, bool negate_odd, float *out,
                              hipStream_t stream);
// phased-array sum at N >= 32768 (fxc_phased.hip); pp: phased_param_bytes()
// per frame of the chunk
size_t phased_param_bytes();
hipError_t launch_phased_params(const SectorArgs &a, int precision, void *pp, int64_t *shifts,
                                hipStream_t stream);
hipError_t launch_phased_large(const SectorArgs &a, int precision, const float2 *z1,
                               const float2 *z2, const void *pp, int n_groups, int accumulate,
                               hipStream_t stream);
hipError_t launch_untangle_large(const float2 *z, int64_t nframes, int M, const float2 *tw_post,
                                 float2 *out, hipStream_t stream);
hipError_t launch_mac_large(const SectorArgs &a, int precision, const float2 *z1, const float2 *z2,
                            const FrameParam *params, int n_groups, int accumulate,
                            hipStream_t stream);
size_t frame_param_bytes();
hipError_t launch_rfft_mfma(const SectorArgs &a, const float *in, int64_t nframes, float2 *out,
                            hipStream_t stream);

// N >= 32768, chosen form: the four-step's first pass with the decode fused
// (fxc_mfma.hip) and its second pass with the untangle and products fused
// (fxc_large.hip).
hipError_t launch_large_pass1_fused(const SectorArgs &a, const int64_t *shifts, float2 *out0,
                                    float2 *out1, int M1, int M2, const float2 *tw_m,
                                    hipStream_t s);
hipError_t launch_reduce_pairs(const double *part, int n_wg, int half, int M1, int M2, double *out,
                               double *scratch, int n_chunks, hipStream_t stream);
hipError_t launch_large_pass2_mac(const SectorArgs &a, const float2 *y1, const float2 *y2,
                                  const FrameParam *params, int M1, int M2, const float2 *tw_m2,
                                  int n_groups, int accumulate, hipStream_t s);
// the phased sum on the fused second pass (N = 16384, M2 = 64); pp: the
// chunk's phased parameters (launch_phased_params); raw: also the first
// half of the inverse FFT into a.cb_out (M per frame), finished by
// launch_irfft_cols_quantise (tw_m: W_M table)
hipError_t launch_large_pass2_phased(const SectorArgs &a, const float2 *y1, const float2 *y2,
                                     const void *pp, int M1, int M2, const float2 *tw_m2,
                                     const float2 *tw_m, bool raw, int n_groups, int accumulate,
                                     hipStream_t s);
hipError_t launch_irfft_cols_quantise(const SectorArgs &a, const float2 *in, int64_t nframes,
                                      uint8_t *out, bool negate_odd, const float2 *tw_m,
                                      hipStream_t s);
hipError_t launch_synth(uint8_t *d_out, int64_t nsamples, int bits, const int32_t *shuffle,
                        int lsb, uint64_t seed, int station, double rho, int64_t delay,
                        hipStream_t stream);
hipError_t launch_synth_spectrum(float2 *out, int64_t first_frame, int64_t nframes, int M,
                                 int station, int64_t shift, uint64_t seed, double rho_c,
                                 double rho_l, double line_bin, double line_fwhm,
                                 hipStream_t stream);
hipError_t launch_synth_pack(const float *x, int64_t nsamples, int bits, int lsb,
                             const int32_t *shuffle, double thresh, double clip, uint8_t *out,
                             hipStream_t stream);

}  // namespace fxc
