// Seeded synthetic station data generated on the device (input generation
// for benchmarks and large tests; not part of the correlator itself).
//
// Sample n of station s: sqrt(rho)*c[n - s*delay] + sqrt(1-rho)*e_s[n] with
// c, e_s unit Gaussians from a counter-based hash (so both stations see the
// same common signal), odd samples negated for an LSB station, quantised
// (2-bit at +-thresh sigma, 1-bit at 0, wider bits uniformly over +-clip
// sigma), packed LSB-first into little-endian words and put through the
// inverse of the station's decode bit map.
#include "fxc_internal.h"

namespace fxc {

__device__ __forceinline__ uint64_t splitmix(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ float hash_gauss(uint64_t seed, uint64_t stream, int64_t n) {
    uint64_t h = splitmix(seed ^ splitmix(stream * 0x100000001b3ull + (uint64_t)n));
    float u1 = ((float)(h >> 40) + 1.0f) * (1.0f / 16777216.0f);
    float u2 = (float)(h & 0xFFFFFFull) * (1.0f / 16777216.0f);
    return sqrtf(-2.0f * logf(u1)) * cosf(6.28318530717958647692f * u2);
}

// The same Gaussian keyed by one 64-bit counter (stream in the top 4 bits).
__device__ __forceinline__ float hash_gauss_key(uint64_t seed, uint64_t key) {
    uint64_t h = splitmix(seed ^ splitmix(key));
    float u1 = ((float)(h >> 40) + 1.0f) * (1.0f / 16777216.0f);
    float u2 = (float)(h & 0xFFFFFFull) * (1.0f / 16777216.0f);
    return sqrtf(-2.0f * logf(u1)) * cosf(6.28318530717958647692f * u2);
}

struct SynthArgs {
    uint32_t *out;
    int64_t nwords;
    int32_t bits;
    int32_t lsb;
    int32_t station;
    int32_t shifts[32];
    uint64_t seed;
    float a, b;  // sqrt(rho), sqrt(1-rho)
    int64_t delay;
    float thresh, clip;
};

// Quantiser of the generators: 2-bit at +-thresh sigma, 1-bit at 0, wider
// bits uniformly over +-clip sigma.
__device__ __forceinline__ uint32_t synth_code(float x, int bits, float thresh, float clip) {
    if (bits == 2) return (x < -thresh) ? 0u : (x < 0.0f) ? 1u : (x < thresh) ? 2u : 3u;
    if (bits == 1) return (x < 0.0f) ? 0u : 1u;
    const int nlev = 1 << bits;
    const float step = 2.0f * clip / (float)nlev;
    float c = floorf(x / step + 0.5f * (float)nlev);
    c = fminf(fmaxf(c, 0.0f), (float)(nlev - 1));
    return (uint32_t)c;
}

// out bit shifts[p] <- in bit p: the inverse of the decode bit map.
__device__ __forceinline__ uint32_t synth_shuffle(uint32_t word, const int32_t *shifts) {
    uint32_t o = 0;
#pragma unroll
    for (int p = 0; p < 32; ++p) o |= ((word >> p) & 1u) << shifts[p];
    return o;
}

__global__ void k_synth(SynthArgs s) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= s.nwords) return;
    const int spw = 32 / s.bits;
    const uint32_t mask = (s.bits >= 32) ? 0xFFFFFFFFu : ((1u << s.bits) - 1u);
    uint32_t word = 0;
    for (int j = 0; j < spw; ++j) {
        const int64_t n = w * spw + j;
        float x = s.a * hash_gauss(s.seed, 0, n - (int64_t)s.station * s.delay) +
                  s.b * hash_gauss(s.seed, 1 + (uint64_t)s.station, n);
        if (s.lsb && (n & 1)) x = -x;
        word |= (synth_code(x, s.bits, s.thresh, s.clip) & mask) << (j * s.bits);
    }
    s.out[w] = synth_shuffle(word, s.shifts);
}

// ------------------------------------------------ spectral-profile generator
// One station's half spectra, frame by frame (frames are independent units
// of the correlator, so a signal built per frame in the frequency domain is
// exactly what each frame's FFT sees before quantisation).  Channel k of
// station 2 carries sky channel k - shift of station 1's grid.  In sky
// channel q the common signal C has coherence rho(q) = rho_c + (rho_l -
// rho_c) exp(-(q - q0)^2 / (2 s^2)) with each station's own noise E:
// X[k] = sqrt(N) (sqrt(rho) C[q] + sqrt(1 - rho) E[k]), C and E complex unit
// Gaussians from the counter hash keyed by (frame, channel), so both
// stations see the same C.  sqrt(N) makes the inverse FFT's samples unit
// variance.
struct SpecArgs {
    float2 *out;         // [frame][M + 1]
    int64_t first_frame, nframes;
    int32_t M, station;
    int64_t shift;
    uint64_t seed;
    float rho_c, rho_l, line_bin, inv_2s2, amp;
};

__global__ void k_synth_spectrum(SpecArgs a) {
    const int half = a.M + 1;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nframes * half) return;
    const int64_t f = i / half;
    const int k = (int)(i - f * half);
    const int64_t q = a.station == 0 ? (int64_t)k : (int64_t)k - a.shift;   // sky channel
    float rho = a.rho_c;
    if (a.rho_l > 0.0f) {
        const float d = (float)q - a.line_bin;
        rho += (a.rho_l - a.rho_c) * expf(-d * d * a.inv_2s2);
    }
    const float g = sqrtf(rho), h = sqrtf(1.0f - rho);
    // counters: (frame < 2^38) << 21 | 2 * (channel + 2^19) | re/im, below
    // 2^60; stream 0 = common signal C, 1 + station = that station's noise
    const uint64_t frame = (uint64_t)(a.first_frame + f);
    const uint64_t ic = (frame << 21) + 2 * (uint64_t)(q + (1 << 19));
    const uint64_t ie = ((2ull + (uint64_t)a.station) << 60) | ((frame << 21) + 2 * (uint64_t)(k + (1 << 19)));
    float re = g * hash_gauss_key(a.seed, ic) + h * hash_gauss_key(a.seed, ie);
    float im = g * hash_gauss_key(a.seed, ic + 1) + h * hash_gauss_key(a.seed, ie + 1);
    if (k == 0 || k == a.M) {
        im = 0.0f;
    } else {
        re *= 0.70710678118654752f;
        im *= 0.70710678118654752f;
    }
    a.out[i] = make_float2(a.amp * re, a.amp * im);
}

hipError_t launch_synth_spectrum(float2 *out, int64_t first_frame, int64_t nframes, int M,
                                 int station, int64_t shift, uint64_t seed, double rho_c,
                                 double rho_l, double line_bin, double line_fwhm,
                                 hipStream_t stream) {
    SpecArgs a;
    a.out = out;
    a.first_frame = first_frame;
    a.nframes = nframes;
    a.M = M;
    a.station = station;
    a.shift = shift;
    a.seed = seed;
    a.rho_c = (float)rho_c;
    a.rho_l = (float)rho_l;
    a.line_bin = (float)line_bin;
    const double sig = line_fwhm / (2.0 * sqrt(2.0 * log(2.0)));
    a.inv_2s2 = sig > 0.0 ? (float)(1.0 / (2.0 * sig * sig)) : 0.0f;
    a.amp = (float)sqrt(2.0 * (double)M);
    const int64_t n = nframes * (int64_t)(M + 1);
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_synth_spectrum, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, a);
    return hipGetLastError();
}

// Unit-variance samples (nwords * 32 / bits of them, frames contiguous) ->
// codes -> LSB-first words through the inverse decode bit map.
struct PackArgs {
    const float *x;
    uint32_t *out;
    int64_t nwords;
    int32_t bits, lsb;
    int32_t shifts[32];
    float thresh, clip;
};

__global__ void k_synth_pack(PackArgs a) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= a.nwords) return;
    const int spw = 32 / a.bits;
    const uint32_t mask = (1u << a.bits) - 1u;
    uint32_t word = 0;
    for (int j = 0; j < spw; ++j) {
        const int64_t n = w * spw + j;
        float x = a.x[n];
        if (a.lsb && (n & 1)) x = -x;
        word |= (synth_code(x, a.bits, a.thresh, a.clip) & mask) << (j * a.bits);
    }
    a.out[w] = synth_shuffle(word, a.shifts);
}

hipError_t launch_synth_pack(const float *x, int64_t nsamples, int bits, int lsb,
                             const int32_t *shuffle, double thresh, double clip, uint8_t *out,
                             hipStream_t stream) {
    PackArgs a;
    a.x = x;
    a.out = reinterpret_cast<uint32_t *>(out);
    a.nwords = nsamples * bits / 32;
    a.bits = bits;
    a.lsb = lsb;
    for (int i = 0; i < 32; ++i) a.shifts[i] = shuffle[i];
    a.thresh = (float)thresh;
    a.clip = (float)clip;
    if (a.nwords <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_synth_pack, dim3((unsigned)((a.nwords + 255) / 256)), dim3(256), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_synth(uint8_t *d_out, int64_t nsamples, int bits, const int32_t *shuffle,
                        int lsb, uint64_t seed, int station, double rho, int64_t delay,
                        hipStream_t stream) {
    SynthArgs s;
    s.out = reinterpret_cast<uint32_t *>(d_out);
    s.nwords = nsamples * bits / 32;
    s.bits = bits;
    s.lsb = lsb;
    s.station = station;
    for (int i = 0; i < 32; ++i) s.shifts[i] = shuffle[i];
    s.seed = seed;
    s.a = (float)sqrt(rho);
    s.b = (float)sqrt(1.0 - rho);
    s.delay = delay;
    s.thresh = 0.98f;
    s.clip = 4.0f;
    const int64_t blocks = (s.nwords + 255) / 256;
    hipLaunchKernelGGL(k_synth, dim3((unsigned)blocks), dim3(256), 0, stream, s);
    return hipGetLastError();
}

}  // namespace fxc
