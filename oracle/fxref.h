/*
 * fxref — CPU oracle for the fx-corr FX hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (libfxcorr.so) never links, loads or calls it.
 *
 * It restates, in plain C, the reference's per-frame algorithm
 * (src/utils.rs, src/main.rs, src/args.rs of M-AKIMOTOO/fx-corr), each function
 * citing the file:line it follows.  Two modes:
 *   FXR_MODE_F32 — reference-faithful: f32 arithmetic in the reference's order
 *                  (built with -ffp-contract=off, as Rust never contracts),
 *                  f32 phase recurrences, f64 accumulators.
 *   FXR_MODE_F64 — FP64 oracle: exact decode, FP64 FFT, per-bin FP64 phases,
 *                  FP64 fringe phasor, FP64 products; in the phased branch
 *                  FP64 rotations and weighted sum, FP64 inverse FFT, the
 *                  quantiser on the samples rounded to f32.
 *
 * The reference's FFT is the third-party crate realfft 3.5.0 over rustfft 6.4.0
 * (not vendored, not buildable here).  fxr_rfft_f32 restates realfft's
 * published method (complex FFT of N/2 points + split post-processing) with an
 * iterative radix-2 FFT; its bit patterns are NOT rustfft's.  Parity of the FFT
 * is pinned only to tolerance (numpy FP64 rfft + closed forms), see tests.
 *
 * Pinning status (see tests/test_oracle_*.py):
 *   parse_shuffle / parse_levels / mean power   pinned (KATs, invariants)
 *   decode (all kinds)                         pinned (brute force, KAT 0xE4)
 *   shift                                      pinned (closed form)
 *   rfft / irfft                               pinned to tolerance (numpy FP64)
 *   delay entry / split / carrier phase        pinned (README KATs, closed forms)
 *   band alignment                             pinned (closed forms)
 *   xcf phase start/step, delay rotation       pinned (closed forms)
 *   sector MAC                                 pinned (invariants, numpy re-run)
 *   phased branch (both modes)                 pinned (closed forms, numpy re-run)
 *   quantise                                   pinned (round trip, nearest level)
 */
#ifndef FXREF_H
#define FXREF_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { FXR_MODE_F32 = 0, FXR_MODE_F64 = 1 };
enum { FXR_GRID_ANT1 = 0, FXR_GRID_ANT2 = 1 };
enum { FXR_SHUFFLE_IDENTITY = 0, FXR_SHUFFLE_PAIRSWAP = 1, FXR_SHUFFLE_GENERIC = 2 };

/* args.rs:202-210.  0 on success, -1 on error. */
int fxr_parse_shuffle(const char *list, int32_t internal_out[32]);
/* args.rs:194-200.  Returns number of levels (== 1<<bits) or -1. */
int fxr_parse_levels(int bits, const char *list, double *out, int cap);
/* main.rs:378-387.  Returns P, or a negative value when invalid. */
double fxr_quantization_mean_power(const double *levels, int n);

typedef struct fxr_decode_plan fxr_decode_plan;
/* utils.rs:173-213 */
int fxr_build_decode_plan(int bits, const int32_t shuffle_internal[32],
                          const double *levels, int nlevels,
                          fxr_decode_plan **out);
void fxr_free_decode_plan(fxr_decode_plan *p);
int fxr_decode_plan_shuffle_kind(const fxr_decode_plan *p);
/* utils.rs:159-171 */
uint32_t fxr_apply_shuffle_word(uint32_t word, const fxr_decode_plan *p);
/* utils.rs:215-376.  0 ok, -1 error (length mismatch / short output). */
int fxr_decode_block(const uint8_t *raw, size_t raw_len, size_t samples,
                     const fxr_decode_plan *p, float *out, size_t out_len,
                     int lsb_to_usb);
/* utils.rs:126-145 */
void fxr_shift_zerofill(float *x, size_t n, int64_t shift);
void fxr_shift_zerofill_f64(double *x, size_t n, int64_t shift);

/* utils.rs:37-59 (realfft r2c, unnormalised).  n: power of two >= 4.
 * out: (n/2+1) interleaved complex. */
int fxr_rfft_f32(int n, const float *in, float *out);
int fxr_rfft_f64(int n, const double *in, double *out);
/* utils.rs:60-66 (realfft c2r, then *= 1/n in f32).  in: (n/2+1) complex. */
int fxr_irfft_f32(int n, const float *in, float *out);
int fxr_irfft_f64(int n, const double *in, double *out);

/* utils.rs:93-124 */
void fxr_apply_delay_and_rate_range(float *spec, size_t len, size_t fft_len,
                                    double df, double delay_s, double rate,
                                    double accel, double t, int is_lsb,
                                    size_t start_bin, size_t end_bin);

/* main.rs:161-182 (DelayEvalConfig; the table is borrowed). */
typedef struct {
    double frame_dt, fs, time_offset_s;
    const double *geom_table; /* NULL when absent */
    int64_t geom_table_len;
    double coarse_delay_s, delay_user_samples, extra_delay_rate_sps;
    double clock1_delay_s, clock1_rate_sps, clock2_delay_s, clock2_rate_sps;
    double d_seek;
    double net_d1_base, total_rate1_base, total_accel1_base;
    double net_d2_base, total_rate2_base, total_accel2_base;
    double lo1_hz, lo2_hz;
} fxr_delay_cfg;

/* main.rs:151-159 */
typedef struct {
    int64_t int1, int2;
    double frac1, frac2;
    float fr_lo1[2], fr_lo2[2];
} fxr_frame_delay;

/* main.rs:184-241 */
void fxr_compute_frame_delay(int64_t frame_idx, const fxr_delay_cfg *cfg,
                             fxr_frame_delay *out);
/* FP64 variant: identical tau/int/frac; fringe phase kept in f64. */
void fxr_compute_frame_delay_f64(int64_t frame_idx, const fxr_delay_cfg *cfg,
                                 fxr_frame_delay *out, double fr1[2],
                                 double fr2[2]);
/* main.rs:126-130 */
void fxr_split_delay(double delay_s, double fs, int64_t *n, double *frac);
/* main.rs:110-120 */
void fxr_carrier_phase(double f_hz, double tau_s, float out[2]);
/* main.rs:133-149 */
void fxr_xcf_phase_start_step(double df, double frac1, double frac2,
                              size_t sb1, size_t sb2, float start[2],
                              float step[2]);
/* Rust `f as i64` (saturating, NaN -> 0) and `f.round()` helpers. */
int64_t fxr_f64_to_i64_sat(double v);

/* main.rs:360-372.  out = {shift_bins, a1s, a1e, a2s}.  0 ok, -1 no overlap. */
int fxr_band_alignment(int fft, double fs, double c1_mhz, double c2_mhz,
                       double bw1_mhz, double bw2_mhz, int64_t out[4]);

/* utils.rs:378-410.  shuffle_out: the internal map of station 1 (sh1).
 * Writes ceil(n*bits/32)*4 bytes; returns that count or -1. */
long fxr_quantise_frame(const float *x, size_t n, int bits,
                        const double *levels, int nlevels,
                        const int32_t shuffle_out[32], uint8_t *out,
                        size_t cap);

/* Per-sector compute of main.rs:2173-2591 (both the accumulate-only branch
 * and the write_raw / phased branch). */
typedef struct {
    int fft_len;
    int bits1, bits2;
    int32_t shuffle1[32], shuffle2[32]; /* internal maps */
    const double *levels1;
    int nlevels1;
    const double *levels2;
    int nlevels2;
    int lsb1, lsb2;
    int64_t a1s, a1e, a2s; /* band alignment */
    int out_grid;          /* FXR_GRID_* */
    double df_hz;
    int need_acf, need_xcf, need_phased, acf_overlap_only, write_raw;
    double w1, w2;
    int bit_out;
    int output_lsb;
    int cpu_threads; /* frames_per_job heuristic, main.rs:2441 */
} fxr_sector_desc;

/* Returns 0 ok, -1 bad arguments.  Accumulators (len n/2+1; acc12 is 2*(n/2+1)
 * interleaved) are OVERWRITTEN.  raw_out (nf*bpf_o bytes) used when
 * write_raw.  threads: worker threads (>=1).  n_failed: sector_failures. */
int fxr_sector_run(const fxr_sector_desc *d, const uint8_t *raw1,
                   const uint8_t *raw2, int64_t first_frame, int64_t nf,
                   const fxr_delay_cfg *cfg, int mode, int threads,
                   double *acc_ph, double *acc11, double *acc12,
                   double *acc22, uint8_t *raw_out, int64_t *n_failed);

#ifdef __cplusplus
}
#endif
#endif
