/*
 * fxcorr.h — C ABI of the MI355X FX hot path (libfxcorr.so).
 *
 * Replaces the per-sector compute of M-AKIMOTOO/fx-corr (src/main.rs:2173-2591)
 * and the per-frame helpers it calls (src/utils.rs, src/main.rs:110-241,
 * src/args.rs:194-210).  The reference is Rust with no FFI; the entry points
 * below are what an FFI crate for that seam (a Rust `extern "C"` block)
 * would bind.  Plain pointers and sizes only.
 *
 * Conventions
 *   - Every function returns an fxc_status (0 = FXC_OK).  No exceptions cross
 *     the ABI.  fxc_last_error() returns a thread-local message for the last
 *     failure on the calling thread.
 *   - Host buffers are caller-owned and borrowed for the duration of a call.
 *   - Device state (tables, scratch) is owned by the plan.  One plan per host
 *     thread at a time; calls on one plan are serialised by the caller.
 *   - The product path has no CPU fallback: without a usable HIP device,
 *     fxc_plan_create fails with FXC_ERR_HIP.
 */
#ifndef FXCORR_H
#define FXCORR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t fxc_status;
enum {
    FXC_OK = 0,
    FXC_ERR_INVALID_ARG = 1,
    FXC_ERR_UNSUPPORTED = 2,
    FXC_ERR_HIP = 3,
    FXC_ERR_NO_OVERLAP = 4,
    FXC_ERR_PARSE = 5,
    FXC_ERR_IO = 6,
    FXC_ERR_NOMEM = 7
};

/* Output grid: which station's bins index the products (main.rs:243-244, 1212). */
enum { FXC_GRID_ANT1 = 0, FXC_GRID_ANT2 = 1 };

/* Product mask (main.rs:2154-2160). */
enum {
    FXC_PRODUCT_ACF = 1u,              /* need_acf_products                     */
    FXC_PRODUCT_XCF = 2u,              /* need_xcf_products                     */
    FXC_PRODUCT_PHASED = 4u,           /* need_phased_products (write_raw path) */
    FXC_PRODUCT_RAW_OUT = 8u,          /* also emit the quantised .raw frames   */
    FXC_PRODUCT_ACF_OVERLAP_ONLY = 16u /* acf_overlap_only (yi-acf)             */
};

/* Arithmetic form of the hot path. */
enum {
    /* FP32 FFT; XCF phase reproduces the reference's f32 start/step
     * (main.rs:133-149, 2513-2527) and f32 fringe phasors; FP64 integration. */
    FXC_PRECISION_PLAIN_FP32 = 0,
    /* FP64 fringe and per-bin fractional-delay phase (reduced mod 2*pi before
     * the cast), FP64 integration; FFT as the plain form unless the plan is
     * eligible for the FP16-MFMA four-step first stage (see fxc_plan_info).
     * Phased plans: each station's rotation and fringe phasor from FP64
     * phases instead of the reference's f32 recurrence; the .raw codes can
     * then differ from the reference's at quantiser decision boundaries. */
    FXC_PRECISION_CHOSEN = 1
};

/* Per-station decode description.  Replaces build_decode_plan's inputs
 * (utils.rs:173) plus the sideband flag (main.rs:1092). */
typedef struct {
    int32_t bits;                 /* 1..16 bits per sample                          */
    int32_t lsb;                  /* 1 = LSB input: negate odd samples (to USB)     */
    int32_t shuffle_internal[32]; /* out bit i <- in bit shuffle_internal[i]        */
    const double *levels;         /* 2^bits levels (args.rs:194-200)                */
    int32_t nlevels;
} fxc_station_desc;

/* Everything run_once captures for the per-sector closure (main.rs:2073-2160). */
typedef struct {
    int32_t fft_len;              /* power of two, 64..65536 (GPU path)             */
    fxc_station_desc st[2];
    int64_t a1s, a1e, a2s;        /* BandAlignment (main.rs:361)                    */
    int32_t out_grid;             /* FXC_GRID_*                                     */
    double df_hz;                 /* fs / fft_len (main.rs:1320)                    */
    uint32_t products;            /* FXC_PRODUCT_* mask                             */
    double w1, w2;                /* phased weights (main.rs:1306-1307)             */
    int32_t bit_out;              /* phased .raw bits = max(bit1, bit2)             */
    int32_t output_lsb;           /* lsb1 && lsb2 (main.rs:1093)                    */
    int32_t precision;            /* FXC_PRECISION_*                                */
} fxc_plan_desc;

/* DelayEvalConfig (main.rs:161-182) as a POD.  geom_table is the 1-s
 * geometric-delay table (main.rs:1629-1644); NULL selects the analytic branch
 * (main.rs:209-230).  The table is copied to the device on each call. */
typedef struct {
    double frame_dt, fs, time_offset_s;
    const double *geom_table;
    int64_t geom_table_len;
    double coarse_delay_s, delay_user_samples, extra_delay_rate_sps;
    double clock1_delay_s, clock1_rate_sps, clock2_delay_s, clock2_rate_sps;
    double d_seek;
    double net_d1_base, total_rate1_base, total_accel1_base;
    double net_d2_base, total_rate2_base, total_accel2_base;
    double lo1_hz, lo2_hz;
} fxc_delay_cfg;

/* FrameDelayEntry (main.rs:151-159). */
typedef struct {
    int64_t int1, int2;
    double frac1, frac2;
    float fr_lo1[2], fr_lo2[2];
} fxc_frame_delay;

/* BandAlignment (main.rs:361). */
typedef struct {
    int64_t shift_bins, a1s, a1e, a2s;
} fxc_band_alignment;

typedef struct fxc_plan fxc_plan;

/* Plan info for reporting. */
typedef struct {
    int32_t fft_len;
    int32_t half;                 /* fft_len/2 + 1                                  */
    int32_t bpf1, bpf2, bpf_out;  /* bytes per station-frame (main.rs:1214-1215)    */
    int32_t decode_ok1, decode_ok2; /* 0: every frame fails decode (utils.rs:372)   */
    int32_t fft_form;             /* 0 = FP32 Stockham, 1 = FP16/FP32-MFMA four-step
                                     (N = 256..2048), 2 = four-step through HBM
                                     (N >= 32768)                                   */
    int32_t threads_per_frame;
    int32_t device_id;
    int32_t variant;              /* launch configuration in use (fxc_tune_report) */
    int32_t grid_divisor;
    int32_t fft_mfma;             /* 1: the FFT's first DFT (32 points) runs as an
                                     FP16 hi/lo MFMA GEMM (chosen form, levels exact
                                     in FP16: fft_form 1, or fft_form 2 when the plan
                                     has no phased-array products)                  */
} fxc_plan_info;

/* ----------------------------------------------------------------------- */
/* Library / device helpers (no reference counterpart: plumbing for hosts   */
/* that hand over device-resident data).                                    */
/* ----------------------------------------------------------------------- */
const char *fxc_version(void);
const char *fxc_strerror(fxc_status s);
const char *fxc_last_error(void);
fxc_status fxc_device_count(int32_t *n);
/* Free and total memory of the current device (hipMemGetInfo). */
fxc_status fxc_device_mem_info(uint64_t *free_bytes, uint64_t *total_bytes);
fxc_status fxc_dev_alloc(size_t bytes, void **dptr);
fxc_status fxc_dev_free(void *dptr);
fxc_status fxc_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream);
fxc_status fxc_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream);
fxc_status fxc_memset_d(void *dst, int value, size_t bytes, void *stream);
/* Device-to-device copy enqueued on `stream` (NULL: the null stream). */
fxc_status fxc_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream);
fxc_status fxc_stream_create(void **stream);
fxc_status fxc_stream_destroy(void *stream);
fxc_status fxc_stream_sync(void *stream);
fxc_status fxc_event_create(void **ev);
fxc_status fxc_event_destroy(void *ev);
fxc_status fxc_event_record(void *ev, void *stream);
fxc_status fxc_event_elapsed_ms(void *ev_start, void *ev_end, float *ms);
/* Work enqueued on `stream` after this call waits for the work before the
 * event's last record (no reference counterpart).  Inside a graph capture
 * it forks the capture onto another stream or joins it back. */
fxc_status fxc_stream_wait_event(void *stream, void *ev);

/* ----------------------------------------------------------------------- */
/* Host-side helpers of the hot-path boundary                               */
/* ----------------------------------------------------------------------- */
/* args.rs:202-210 parse_shuffle: external list -> internal map. */
fxc_status fxc_parse_shuffle(const char *list, int32_t internal_out[32]);
/* args.rs:194-200 parse_levels.  *n_out = 1<<bits on success. */
fxc_status fxc_parse_levels(int32_t bits, const char *list, double *out,
                            int32_t cap, int32_t *n_out);
/* main.rs:378-387 quantization_mean_power. */
fxc_status fxc_quantization_mean_power(const double *levels, int32_t n,
                                       double *p_out);
/* main.rs:360-372 compute_band_alignment (centres / bandwidths in MHz). */
fxc_status fxc_compute_band_alignment(int32_t fft_len, double fs_hz,
                                      double c1_mhz, double c2_mhz,
                                      double bw1_mhz, double bw2_mhz,
                                      fxc_band_alignment *out);
/* main.rs:2058-2072 sector partition: counts[] (cap entries) of frames per
 * sector; *n_sectors_out. */
fxc_status fxc_sector_partition(int64_t total_frames, double frame_sec,
                                int64_t *counts, int64_t cap,
                                int64_t *n_sectors_out);
/* main.rs:126-130 split_delay_to_integer_and_fractional (host). */
fxc_status fxc_split_delay(double delay_s, double fs_hz, int64_t *n_out,
                           double *frac_out);

/* ----------------------------------------------------------------------- */
/* Plans and the per-sector hot path                                        */
/* ----------------------------------------------------------------------- */
/* build_decode_plan x2 + FftHelper::new + captured constants
 * (utils.rs:173-213, utils.rs:22-29, main.rs:2073-2074). */
fxc_status fxc_plan_create(const fxc_plan_desc *desc, fxc_plan **out);
fxc_status fxc_plan_destroy(fxc_plan *plan);
fxc_status fxc_plan_get_info(const fxc_plan *plan, fxc_plan_info *info);

/* The per-sector closure of run_once (main.rs:2173-2591) on host buffers.
 * raw1/raw2: nf frames of bpf1/bpf2 bytes (len1/len2 >= nf*bpf).
 * first_frame: global frame index of the sector's first frame (emitted).
 * Outputs (length half, acc12 = 2*half interleaved re,im), OVERWRITTEN:
 *   acc11/acc22/acc12 for ACF/XCF, acc_ph for PHASED; any may be NULL when
 *   not requested.  raw_out: nf*bpf_out bytes when FXC_PRODUCT_RAW_OUT.
 * n_failed: frames skipped (sector_failures, main.rs:2221, 2592). */
fxc_status fxc_sector_run(fxc_plan *plan, const uint8_t *raw1, size_t len1,
                          const uint8_t *raw2, size_t len2,
                          int64_t first_frame, int64_t nf,
                          const fxc_delay_cfg *delay, double *acc_ph,
                          double *acc11, double *acc12, double *acc22,
                          uint8_t *raw_out, int64_t *n_failed);

/* Device-resident variant: d_raw1/d_raw2/d_acc/d_raw_out are device pointers.
 * d_acc receives 5*half doubles laid out [ph | 11 | 22 | 12re | 12im].
 * Enqueued on `stream` (NULL = the HIP null stream); returns without waiting.
 * The plan's partial-sum scratch is reused by every call: calls on one plan
 * must be ordered (same stream, or synchronised by the caller). */
fxc_status fxc_sector_run_device(fxc_plan *plan, const uint8_t *d_raw1,
                                 const uint8_t *d_raw2, int64_t first_frame,
                                 int64_t nf, const fxc_delay_cfg *delay,
                                 double *d_acc, uint8_t *d_raw_out,
                                 void *stream);

/* Allocation-free variant of fxc_sector_run_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The call allocates and frees nothing and never synchronises;
 * its scratch (geometric table, workgroup partial sums, reduction sums) is
 * one static device array of FXC_NOALLOC_SCRATCH_BYTES per device, held
 * until the process ends.  Same arguments and d_acc layout as
 * fxc_sector_run_device, for ACF/XCF plans of fft_len <= 16384 whose XCF
 * needs no phase-chain table (the chosen form, or the plain form with at
 * most 512 overlap bins); other plans fail with FXC_ERR_UNSUPPORTED, a
 * geometric table longer than FXC_NOALLOC_MAX_TABLE entries with
 * FXC_ERR_INVALID_ARG.  The table is copied from host memory on each call,
 * so a call with a table cannot be captured into a graph.
 * Every call on one device shares that device's array: calls that overlap
 * on one device (other streams, other host threads) race and must be
 * serialised by the caller.  The plan's current device must be current. */
#define FXC_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
#define FXC_NOALLOC_MAX_TABLE 8192
fxc_status fxc_sector_run_device_noalloc(fxc_plan *plan, const uint8_t *d_raw1,
                                         const uint8_t *d_raw2, int64_t first_frame,
                                         int64_t nf, const fxc_delay_cfg *delay,
                                         double *d_acc, void *stream);

/* ----------------------------------------------------------------------- */
/* Whole-scan cross spectrum (the reference's --fringe XCF pipeline,        */
/* main.rs:1692-1947, replacing the per-chunk rayon map/reduce at           */
/* 1760-1849 and the accumulation at 1850).  `ac` is half complex FP64      */
/* values, re/im interleaved, zeroed by the caller before the first chunk;  */
/* each call adds the chunk's XCF sum (the plan's output grid, frames that  */
/* fail to decode skipped) times 1/N^2.  The plan must request XCF and no   */
/* phased products.  The device variant is enqueued on `stream`.           */
/* ----------------------------------------------------------------------- */
fxc_status fxc_fringe_accumulate(fxc_plan *plan, const uint8_t *raw1, size_t len1,
                                 const uint8_t *raw2, size_t len2, int64_t first_frame,
                                 int64_t nf, const fxc_delay_cfg *delay, double *ac,
                                 int64_t *n_failed);
fxc_status fxc_fringe_accumulate_device(fxc_plan *plan, const uint8_t *d_raw1,
                                        const uint8_t *d_raw2, int64_t first_frame,
                                        int64_t nf, const fxc_delay_cfg *delay,
                                        double *d_ac, void *stream);

/* finalize_cross_spectrum (xcf.rs:13-104) without its plots: zeroes bins 0
 * and fft_len/2 of `spec` (len complex values, re/im interleaved) in place,
 * then returns per bin the magnitude (bin 0: the mean of the others halved),
 * the phase (safe_arg, utils.rs:425), the phase unwrapped from bin 1 on
 * (utils.rs:412-424) and the least-squares line through it against sky
 * frequency (USB: sky_low + k df; LSB: sky_low + bw - k df) over bins
 * 1 .. len-2 (even fft_len), whose slope gives delay = -slope / 2 pi.
 * Output arrays may be NULL. */
typedef struct {
    int32_t has_delay;            /* 0 when the fit is degenerate (xcf.rs:88)  */
    double delay_seconds;
    double slope_rad_per_hz;
    double intercept_rad;
} fxc_cross_spectrum_result;
fxc_status fxc_finalize_cross_spectrum(double *spec, int64_t len, int32_t fft_len,
                                       double sampling_rate_hz, double sky_low_mhz,
                                       int32_t is_lsb, fxc_cross_spectrum_result *res,
                                       double *mag_out, double *phase_out,
                                       double *unwrapped_out, double *fit_out);

/* finalize_auto_spectrum (acf.rs:3-23): out[i] = acc[i] / power_norm (acc[i]
 * itself when power_norm <= 0), then out[0] = 0.  Host function. */
fxc_status fxc_finalize_auto_spectrum(const double *acc, int64_t len, double power_norm,
                                      double *out);

/* The numbers behind yi-phasedarray's plots (main.rs:2727-2768), host
 * function: the phased, station-1 and station-2 auto spectra of a scan
 * finalised with power_norm = max(emitted_frames * fft_len, 1); the first
 * spec_bins bins (fft_len/2 for even fft_len, the whole half spectrum for
 * odd) on the axis obs_mhz + i df; the phased power and the three
 * amplitudes sqrt(power); and the phased autocorrelation: the f32 spectrum
 * mirrored into fft_len bins, inverse DFT x 1/fft_len, real part, rotated
 * so that acf_shifted[fft_len/2] is lag 0 (lags -fft_len/2 ..).  The caller
 * provides arrays of fft_len/2+1 (acf_shifted: fft_len) doubles. */
typedef struct {
    double *freqs_mhz, *phased_power, *amp_ph, *amp11, *amp22;
    double *acf_shifted;
    int64_t spec_bins;            /* set by the call */
} fxc_phased_plot;
fxc_status fxc_phased_plot_series(const double *acc_ph, const double *acc11, const double *acc22,
                                  int64_t len, int32_t fft_len, int64_t emitted_frames,
                                  double obs_mhz, double df_hz, fxc_phased_plot *out);

/* ----------------------------------------------------------------------- */
/* Streaming host input (replaces the reader thread + per-sector loop of    */
/* main.rs:2098-2213 for callers whose sector bytes are in host memory).   */
/* fxc_host_alloc returns page-locked memory: a reader filling such buffers */
/* lets the copies run at PCIe rate.  A pipeline owns `depth` device slots  */
/* of max_frames frames, a copy stream and a compute stream: submit copies  */
/* a sector into the next slot (after that slot's previous sector has read  */
/* it), runs it on the compute stream behind the sectors submitted before,  */
/* and copies its spectra (5 x half f64, the fxc_sector_run_device layout)  */
/* into acc_out; it returns at once with a ticket.  raw1/raw2 must stay     */
/* unchanged, and acc_out is only valid, once fxc_pipeline_wait(ticket)     */
/* has returned.  One pipeline per plan; plans without .raw output.        */
/* ----------------------------------------------------------------------- */
fxc_status fxc_host_alloc(size_t bytes, void **out);
fxc_status fxc_host_free(void *p);
typedef struct fxc_pipeline fxc_pipeline;
fxc_status fxc_pipeline_create(fxc_plan *plan, int64_t max_frames, int32_t depth,
                               fxc_pipeline **out);
fxc_status fxc_pipeline_submit(fxc_pipeline *pl, const uint8_t *raw1, const uint8_t *raw2,
                               int64_t first_frame, int64_t nf, const fxc_delay_cfg *delay,
                               double *acc_out, int64_t *ticket);
fxc_status fxc_pipeline_wait(fxc_pipeline *pl, int64_t ticket);
fxc_status fxc_pipeline_destroy(fxc_pipeline *pl);

/* Kernel timing of the dominant (fused per-frame) kernel measured with HIP
 * events on the launch stream; enabled per plan. */
fxc_status fxc_plan_set_timing(fxc_plan *plan, int32_t enable);
fxc_status fxc_plan_get_timing(fxc_plan *plan, double *total_ms,
                               int64_t *launches);

/* ----------------------------------------------------------------------- */
/* Per-function entry points (each runs the device code of the fused path  */
/* on its own; host buffers).                                               */
/* ----------------------------------------------------------------------- */
/* decode_block_into_with_plan (utils.rs:215-376) + optional
 * apply_integer_sample_shift_zerofill (utils.rs:126-145) for nframes frames
 * of station `station` (0/1).  shifts may be NULL (no shift).
 * out: nframes*fft_len floats.  *n_failed: frames that fail decode. */
fxc_status fxc_decode_frames(fxc_plan *plan, int32_t station,
                             const uint8_t *raw, size_t len, int64_t nframes,
                             const int64_t *shifts, float *out,
                             int64_t *n_failed);
/* FftHelper::forward_r2c_process (utils.rs:37-59) on nframes real frames:
 * in nframes*fft_len floats, out nframes*(half) interleaved complex.  When
 * fxc_plan_info.fft_mfma is 1 the samples enter the FFT as FP16: exact for
 * the plan's levels (what the decode feeds the FFT), rounded otherwise. */
fxc_status fxc_rfft_frames(fxc_plan *plan, const float *in, int64_t nframes,
                           float *out);
/* FftHelper::inverse_c2r_process (utils.rs:60-66): in nframes*half complex,
 * out nframes*fft_len floats (scaled 1/fft_len). */
fxc_status fxc_irfft_frames(fxc_plan *plan, const float *in, int64_t nframes,
                            float *out);
/* compute_frame_delay_entry (main.rs:184-241) evaluated on the device in
 * FP64 for frames first..first+n-1. */
fxc_status fxc_delay_eval(const fxc_delay_cfg *cfg, int64_t first, int64_t n,
                          fxc_frame_delay *out);
/* quantise_frame (utils.rs:378-410) with the plan's bit_out / station-1
 * levels / station-1 shuffle: in nframes*fft_len floats -> nframes*bpf_out. */
fxc_status fxc_quantise_frames(fxc_plan *plan, const float *in, int64_t nframes,
                               uint8_t *out);

/* ----------------------------------------------------------------------- */
/* Synthetic input (no reference counterpart: seeded station data generated */
/* on the device for benchmarks and large tests).                           */
/* ----------------------------------------------------------------------- */
/* Station `station` (0/1) of a correlated pair: unit Gaussians mixed with
 * coefficient rho, station 1 delayed by delay_samples, odd samples negated
 * when lsb, quantised (2-bit at +-0.98 sigma), packed LSB-first into
 * little-endian words and stored through the inverse of shuffle_internal so
 * that decoding with that map recovers the codes.  bits must divide 32 and
 * nsamples*bits must be a multiple of 32.  Enqueued on `stream`. */
fxc_status fxc_synth_station_device(uint8_t *d_out, int64_t nsamples, int32_t bits,
                                    const int32_t shuffle_internal[32], int32_t lsb,
                                    uint64_t seed, int32_t station, double rho,
                                    int64_t delay_samples, void *stream);

/* A seeded two-station pair with a spectral coherence profile, built frame
 * by frame in the frequency domain (frames are the correlator's independent
 * units) and inverse-FFT'd on the device: in sky channel q (station-1 grid)
 * the stations share a common signal with coherence
 *   rho(q) = rho_continuum + (rho_line - rho_continuum) exp(-(q - line_bin)^2 / 2s^2),
 * s = line_fwhm_bins / 2.3548 (no line when rho_line <= 0); station 2's
 * channel k carries sky channel k - shift_bins (the band alignment's
 * a2s - a1s).  Unit-variance samples are quantised per station (1-bit at 0,
 * 2-bit at +-thresh_sigma, wider bits uniformly over +-clip_sigma with the
 * levels l_c = c - (2^bits - 1)/2), odd samples negated when lsb, and packed
 * as fxc_synth_station_device packs.  Writes frames first_frame ..
 * first_frame + nframes - 1 (nframes * fft_len * bits / 8 bytes per
 * station); the same frame always gets the same samples.  Allocates its
 * FFT scratch and synchronises `stream` before returning. */
typedef struct {
    int32_t fft_len;
    int32_t bits[2];
    int32_t lsb[2];
    int32_t shuffle_internal[2][32];
    int64_t shift_bins;
    double rho_continuum, rho_line, line_bin, line_fwhm_bins;
    double thresh_sigma, clip_sigma;
    uint64_t seed;
} fxc_synth_spectral_desc;
fxc_status fxc_synth_spectral_pair_device(const fxc_synth_spectral_desc *desc, int64_t first_frame,
                                          int64_t nframes, uint8_t *d_raw1, uint8_t *d_raw2,
                                          void *stream);

/* Device memory held by a plan (tables + scratch), bytes. */
fxc_status fxc_plan_device_bytes(const fxc_plan *plan, uint64_t *bytes);

/* Autotuner (no reference counterpart).  Times the plan's candidate launch
 * configurations on nf frames of the caller's device data (median of three
 * runs each, HIP events on `stream`) and keeps the fastest in the plan for
 * later calls.  Candidates: the FFT pass sequence (radix-8 or radix-16
 * passes, N = 2048..8192) and workgroups per launch (frames per workgroup)
 * for the Stockham form; the four-step split M1 x M2 and the chunk length
 * for N >= 32768 (the fused four-step, also at N = 16384: its default chunk
 * keeps both stations' intermediate rows within 128 MiB, twice and half
 * that are tried too); the kernel organization for the MFMA form at N = 1024
 * (every wave carrying whole frames, producer/consumer waves, or
 * producer/consumer waves with FP32 batch sums; N = 2048 has the
 * producer/consumer organization only).  Results are the same sums up to
 * FP64 summation order, or FP32 rounding of the batch sums (about 1e-7 of
 * the spectrum's scale) for the FP32-batch organization. */
typedef struct {
    int32_t candidates;
    int32_t variant;           /* Stockham: 0 default passes, 1 radix-16 passes
                                  (N = 2048..8192) or, chosen form at N = 16384, the
                                  fused four-step (split_m1 = 128, the default there;
                                  also for phased plans through fxc_plan_set_tuning,
                                  0 then selects the Stockham phased kernel);
                                  MFMA: 1 producer/consumer (default), 2 producer/consumer
                                  with FP32 sums over batches of 128 frames added to FP64
                                  totals, 0 whole-frame waves (2 and 0: N = 1024 only);
                                  N >= 32768: 1 fused four-step, 0 through HBM */
    int32_t grid_divisor;      /* workgroups = default / grid_divisor */
    int32_t split_m1;          /* four-step forms: column DFT length M1 (else 0) */
    int64_t chunk_frames;      /* four-step forms: frames per chunk (else 0) */
    double best_ms, default_ms;
} fxc_tune_report;
fxc_status fxc_plan_autotune(fxc_plan *plan, const uint8_t *d_raw1, const uint8_t *d_raw2,
                             int64_t nf, const fxc_delay_cfg *delay, void *stream,
                             fxc_tune_report *report);
/* Applies a configuration (variant, grid_divisor, split_m1, chunk_frames of
 * `cfg`; other fields ignored), e.g. one saved from an earlier autotune. */
fxc_status fxc_plan_set_tuning(fxc_plan *plan, const fxc_tune_report *cfg);

/* HIP-graph capture of device-entry work (no reference counterpart: replays
 * a fixed launch sequence, e.g. the sectors of a scan, without per-launch
 * host cost).  Between begin and end, only fxc_sector_run_device calls on
 * `stream` (non-NULL) may be made, with plan timing off and each plan's
 * geometric table already uploaded by an earlier call with the same table;
 * the captured work never synchronises or clears memory with a memset node.
 * Replays read and write the same device buffers as the captured calls:
 * the caller's inputs and outputs, and the scratch of every plan whose work
 * was captured.  While a graph that captured a plan is alive:
 *   - a call on that plan that would grow its scratch, or that brings a
 *     different geometric table, fails with FXC_ERR_INVALID_ARG before any
 *     launch (size the scratch with the largest call before capturing);
 *   - fxc_plan_destroy returns at once and the plan's memory is released
 *     when the last such graph is destroyed.
 * A call that would grow plan scratch inside a capture fails the same way.
 * The capture may fork onto other streams and join back through events
 * (fxc_event_record on the capturing stream, fxc_stream_wait_event on the
 * other): sector calls on those streams join the graph as parallel
 * branches, and their plans are held like the capturing stream's.  End the
 * capture on the stream that began it, after every branch has joined. */
typedef struct fxc_graph fxc_graph;
fxc_status fxc_graph_begin(void *stream);
fxc_status fxc_graph_end(void *stream, fxc_graph **out);
fxc_status fxc_graph_launch(fxc_graph *g, void *stream);
fxc_status fxc_graph_destroy(fxc_graph *g);
/* Live graphs holding the plan's buffers. */
fxc_status fxc_plan_live_graphs(const fxc_plan *plan, int32_t *n);

/* Device error word of a plan (no reference counterpart).  The fused MFMA
 * kernels bound every wait between their producer and consumer waves; a
 * workgroup whose wait runs out writes NaN partial sums and sets the plan's
 * error word.  fxc_sector_run, fxc_pipeline_wait and the server's job
 * completion read it and fail with FXC_ERR_HIP; callers of
 * fxc_sector_run_device read and clear it with this call once their work
 * on `stream` is done (it synchronises `stream`). */
fxc_status fxc_plan_check_device_error(fxc_plan *plan, void *stream);

/* ----------------------------------------------------------------------- */
/* Serving (no reference counterpart: the reference runs one scan per       */
/* process; this admits correlation jobs continuously).  Jobs run the same  */
/* sector path as fxc_sector_run_device; each job's input arrives in chunks */
/* (frames in order, e.g. 64 MB e-VLBI blocks), its delay-model frame index */
/* and FP64 partial spectra stay resident on the device between chunks, and */
/* jobs are admitted while others run as slots and HBM budget free up.      */
/* Host input is copied on a copy stream per job stream into two windows    */
/* per station, alternating between chunks.  Device buffers, page-locked    */
/* result buffers and events of finished jobs are pooled for later jobs     */
/* (pooled bytes count against hbm_budget_bytes and are released, largest   */
/* first, when an admission needs the room, and at fxc_server_destroy).     */
/* ----------------------------------------------------------------------- */
typedef struct fxc_server fxc_server;
typedef struct {
    uint64_t hbm_budget_bytes;  /* admission budget for plans + input windows + spectra */
    int32_t max_active_jobs;    /* concurrent job slots */
    int32_t n_streams;          /* HIP streams the jobs are spread over (1..4) */
} fxc_server_cfg;
typedef struct {
    fxc_plan_desc plan;         /* levels pointers are copied at submit */
    fxc_delay_cfg delay;        /* geometric table copied at submit */
    int64_t first_frame;        /* delay-model index of the job's frame 0 */
    int32_t input_on_device;    /* chunks point to device (1) or host (0) memory */
} fxc_job_desc;
typedef struct {
    double queue_ms;            /* submit -> admission */
    double latency_ms;          /* submit -> result on the host */
    int64_t frames;
    uint64_t footprint_bytes;   /* device bytes held while active */
} fxc_job_stats;
typedef struct {
    int64_t jobs_done, jobs_failed;
    double p50_latency_ms, p99_latency_ms, mean_latency_ms;
    int32_t max_concurrent_jobs;
    uint64_t peak_hbm_bytes;
} fxc_server_stats;

fxc_status fxc_server_create(const fxc_server_cfg *cfg, fxc_server **out);
/* Finishes the admitted and queued work, then frees everything. */
fxc_status fxc_server_destroy(fxc_server *s);
fxc_status fxc_job_submit(fxc_server *s, const fxc_job_desc *job, uint64_t *job_id);
/* Append nframes frames (raw1: nframes*bpf1 bytes, raw2: nframes*bpf2); the
 * buffers must stay valid until the job completes. */
fxc_status fxc_job_push(fxc_server *s, uint64_t job_id, const uint8_t *raw1,
                        const uint8_t *raw2, int64_t nframes);
/* No more input: the job completes after its last chunk. */
fxc_status fxc_job_close(fxc_server *s, uint64_t job_id);
/* Waits up to timeout_s (< 0: forever); *done = 1 when the result is ready. */
fxc_status fxc_job_wait(fxc_server *s, uint64_t job_id, double timeout_s, int32_t *done);
/* [ph | 11 | 22 | 12re | 12im] x (fft_len/2+1) FP64 sums over all chunks. */
fxc_status fxc_job_result(fxc_server *s, uint64_t job_id, double *acc, fxc_job_stats *stats);
fxc_status fxc_job_release(fxc_server *s, uint64_t job_id);
fxc_status fxc_server_get_stats(fxc_server *s, fxc_server_stats *out);

/* ----------------------------------------------------------------------- */
/* Input: packed-sample reader (replaces PackedSampleReader, main.rs:458-516, */
/* and the word-aligned seek of main.rs:1578-1616).                          */
/* ----------------------------------------------------------------------- */
typedef struct fxc_reader fxc_reader;

/* PackedSampleReader::open (main.rs:465-482): seek to start_byte; with a
 * non-zero bit_offset (0..7) the first byte is held back for re-alignment. */
fxc_status fxc_reader_open(const char *path, uint64_t start_byte, int32_t bit_offset,
                           fxc_reader **out);
/* read_packed_with_padding (main.rs:484-499, read_with_padding 445-456):
 * fills len bytes, zero-padded past end of file; *got = bytes actually read
 * from the file (may be NULL). */
fxc_status fxc_reader_read(fxc_reader *r, uint8_t *out, size_t len, size_t *got);
fxc_status fxc_reader_close(fxc_reader *r);
/* main.rs:1585-1600: start sample floored to whole 32-bit words of `bits`
 * samples; start byte = actual*bits/8; residual = actual - requested (<= 0). */
fxc_status fxc_word_aligned_start(uint64_t start_samples, int32_t bits,
                                  uint64_t *actual_samples, uint64_t *start_byte,
                                  int64_t *residual_samples);

/* ----------------------------------------------------------------------- */
/* Output: sector normalisation (main.rs:2645-2683) and the .cor writer     */
/* (CorWriter, cor.rs:152-296).                                             */
/* ----------------------------------------------------------------------- */
/* acc in the [ph | 11 | 22 | 12re | 12im] x (fft_len/2+1) layout of
 * fxc_sector_run_device; each output (may be NULL) gets fft_len/2 complex
 * f32 (re, im) values: ph / (0.5 max(w1^2 P1 + w2^2 P2, 1e-12) nf N^2),
 * 11 / (0.5 P1 nf N^2), 22 / (0.5 P2 nf N^2), 12 / (0.5 sqrt(P1 P2) nf N^2). */
fxc_status fxc_normalize_sector(const double *acc, int32_t fft_len, int64_t nf,
                                double level_power1, double level_power2, double w1, double w2,
                                float *out_ph, float *out11, float *out22, float *out12);

typedef struct {            /* CorStation (cor.rs:26-31) */
    const char *name;       /* first 16 bytes used */
    uint8_t code;
    double ecef_m[3];
} fxc_cor_station;
typedef struct {            /* CorHeaderConfig (cor.rs:33-43) */
    int32_t sampling_speed_hz;
    double observing_frequency_hz;
    int32_t fft_point;
    int32_t number_of_sector_hint;
    int64_t clock_reference_unix_sec;
    const char *source_name;
    double source_ra_rad, source_dec_rad;
} fxc_cor_header;
typedef struct {            /* CorClockModel (cor.rs:45-54) */
    uint32_t sec, nsec;
    double delay, rate, acel, jerk, snap;
} fxc_cor_clock_model;
typedef struct {            /* CorSectorModel (cor.rs:56-62) */
    fxc_cor_clock_model station1, station2;
    float amp[2];
    int16_t phs[2];
} fxc_cor_sector_model;
typedef struct fxc_cor_writer fxc_cor_writer;

/* build_file_header (cor.rs:153-203): the 256-byte file header. */
fxc_status fxc_cor_build_header(const fxc_cor_header *cfg, const fxc_cor_station *st1,
                                const fxc_cor_station *st2, uint8_t out[256]);
/* The 128-byte sector header of write_sector_with_model (cor.rs:243-276);
 * model may be NULL. */
fxc_status fxc_cor_build_sector_header(int64_t unix_sec, float integ_s,
                                       const fxc_cor_sector_model *model, uint8_t out[128]);
/* CorWriter::create / write_sector_with_model / finalize (cor.rs:206-296):
 * spectrum = fft_point/2 complex f32 (re, im) little-endian. */
fxc_status fxc_cor_create(const char *path, const fxc_cor_header *cfg,
                          const fxc_cor_station *st1, const fxc_cor_station *st2,
                          fxc_cor_writer **out);
fxc_status fxc_cor_write_sector(fxc_cor_writer *w, int64_t unix_sec, float integ_s,
                                const float *spectrum, int64_t n_complex,
                                const fxc_cor_sector_model *model);
fxc_status fxc_cor_finalize(fxc_cor_writer *w);

/* ----------------------------------------------------------------------- */
/* Station geometry (geom.rs) and the 1-second geometric-delay table of    */
/* the delay model (main.rs:1628-1642).  Host functions.  Sidereal time is */
/* IAU 2006 GMST plus the equation of the equinoxes with the IAU 2000B     */
/* nutation series (within 1 mas of the 2000A series behind ERFA's         */
/* eraGst06a); precession is ERFA's eraPmat06 model (IAU 2006 with frame   */
/* bias).  MJDs stand for both UT1 and TT, as in geom.rs:282-285.          */
/* ----------------------------------------------------------------------- */
typedef struct {
    double delay1_s, delay2_s;   /* -(x_i . s)/c per antenna at the epoch           */
    double delay_s;              /* baseline delay -((x2 - x1) . s)/c                */
    double rate_sps;             /* central difference over +-1 s                    */
    double accel_sps2;           /* second difference over +-1 s                     */
} fxc_geometric_delay_result;

/* parse_ra (geom.rs:86-111): "16h42m58.8s", "16:42:58.8", packed 164258.8
 * (|v| >= 10000) are hours; a plain number below 10000 is degrees. */
fxc_status fxc_parse_ra(const char *text, double *radians);
/* parse_dec (geom.rs:115-148): "+39d48m36.0s", "39:48:36", packed 394836.0,
 * or decimal degrees; a leading '-' negates. */
fxc_status fxc_parse_dec(const char *text, double *radians);
/* parse_epoch_to_mjd (geom.rs:226-257): ISO "2024-02-12T15:52:00Z",
 * "YYYY/DDD [HH:MM:SS]", an MJD in [40000, 100000) or a Julian year in
 * [1800, 3000) (optionally prefixed J). */
fxc_status fxc_parse_epoch_mjd(const char *text, double *mjd);
/* precess_j2000_to_mean_of_date (geom.rs:265-277). */
fxc_status fxc_precess_j2000_to_date(double ra, double dec, double mjd, double *ra_date,
                                     double *dec_date);
/* mjd_to_gmst (geom.rs:279-287) and the apparent sidereal time of
 * source_vector_itrs (geom.rs:289-297), radians in [0, 2 pi); either
 * output may be NULL. */
fxc_status fxc_sidereal_time(double mjd, double *gmst, double *gast);
/* calculate_geometric_delay_and_derivatives (geom.rs:346-384): ra/dec of
 * date (radians), ECEF positions in metres. */
fxc_status fxc_geometric_delay(const double ant1_ecef[3], const double ant2_ecef[3], double ra,
                               double dec, double mjd, fxc_geometric_delay_result *out);
/* The 1-second table (main.rs:1628-1642): table[k] = baseline delay at
 * mjd + (time_offset_s + k)/86400 for k < n_points; the delay model reads
 * it as fxc_delay_cfg.geom_table. */
fxc_status fxc_geometric_delay_table(const double ant1_ecef[3], const double ant2_ecef[3],
                                     double ra, double dec, double mjd, double time_offset_s,
                                     int64_t n_points, double *table);
/* Table length for a process of total_frames frames of frame_dt seconds:
 * max(ceil(duration), 1) + 1 (main.rs:1630-1634). */
int64_t fxc_geometric_table_points(int64_t total_frames, double frame_dt);
/* How far this library's geometric delays may move the fringe phase from
 * the reference's (no reference counterpart: a check for callers on long
 * baselines).  Sidereal time here uses the IAU 2000B nutation series where
 * the reference's eraGst06a uses 2000A; 2000B is within 1 mas of 2000A over
 * 1995-2050 (its published accuracy), so the hour angle differs by at most
 * 4.85e-9 rad and the delay by |ant2 - ant1| * 4.85e-9 / c.  Returns that
 * delay bound (s) and the fringe-phase bound 2 pi obs_hz delay (rad); the
 * plain form's 1e-5 relative .cor bar holds end to end for baselines whose
 * phase bound is below 1e-5 (about 15 m at 6.7 GHz; 7.4e-5 rad at 108 m). */
fxc_status fxc_geometric_model_bound(const double ant1_ecef[3], const double ant2_ecef[3],
                                     double obs_hz, double *delay_s, double *phase_rad);

#ifdef __cplusplus
}
#endif
#endif /* FXCORR_H */
