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

__global__ void k_synth(SynthArgs s) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= s.nwords) return;
    const int spw = 32 / s.bits;
    const uint32_t mask = (s.bits >= 32) ? 0xFFFFFFFFu : ((1u << s.bits) - 1u);
    const int nlev = 1 << s.bits;
    uint32_t word = 0;
    for (int j = 0; j < spw; ++j) {
        const int64_t n = w * spw + j;
        float x = s.a * hash_gauss(s.seed, 0, n - (int64_t)s.station * s.delay) +
                  s.b * hash_gauss(s.seed, 1 + (uint64_t)s.station, n);
        if (s.lsb && (n & 1)) x = -x;
        uint32_t code;
        if (s.bits == 2) code = (x < -s.thresh) ? 0u : (x < 0.0f) ? 1u : (x < s.thresh) ? 2u : 3u;
        else if (s.bits == 1) code = (x < 0.0f) ? 0u : 1u;
        else {
            float step = 2.0f * s.clip / (float)nlev;
            float c = floorf(x / step + 0.5f * (float)nlev);
            c = fminf(fmaxf(c, 0.0f), (float)(nlev - 1));
            code = (uint32_t)c;
        }
        word |= (code & mask) << (j * s.bits);
    }
    uint32_t o = 0;
#pragma unroll
    for (int p = 0; p < 32; ++p) o |= ((word >> p) & 1u) << s.shifts[p];
    s.out[w] = o;
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
