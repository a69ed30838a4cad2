// The reference's f32 phase chains, replayed per frame for the plain form.
//
// XCF phase (main.rs:2513-2527, start/step from xcf_phase_start_and_step,
// main.rs:133-149):
//     p_0 = from_polar(1, ang0 as f32),  p_{k+1} = p_k * from_polar(1, angs as f32)
// Phased per-station rotation (utils.rs:93-117 with start_bin = 0, rate =
// accel = t = 0, USB sign; main.rs:2238-2262):
//     r_0 = from_polar(1, (step * 0) as f32),  r_{b+1} = r_b * from_polar(1, step as f32)
//
// Each link is Rust's Complex<f32> product: four f32 products and two f32
// sums, no fused multiply-add, so a chain carries the reference's rounding
// bit for bit.  from_polar's cos/sin are the correctly rounded f32 values;
// for the step angles of N >= 2048 (|angle| < 3.1e-3) glibc's cosf/sinf
// return exactly those, and a start one ulp away (larger band-shifted start
// angles) moves a chain by at most ~5e-7 over 30720 links.
//
// One thread per chain writes its row in pairs of links (16-byte stores).
// Blocks of one wave, so that the few long chains of a launch spread over
// the compute units (each chain's walk is latency-bound).
#include "fxc_dev.h"

namespace fxc {

__device__ __forceinline__ float2 c_mul_rust(float2 a, float2 b) {
#pragma clang fp contract(off)
    const float rr = a.x * b.x;
    const float ii = a.y * b.y;
    const float ri = a.x * b.y;
    const float ir = a.y * b.x;
    return make_float2(rr - ii, ri + ir);
}

__global__ __launch_bounds__(64) void k_phase_chain(SectorArgs a, int kind, float2 *out) {
    const int nch = kind == CHAIN_XCF ? 1 : 2;
    const int len = kind == CHAIN_XCF ? a.ov : a.M + 1;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nf * nch) return;
    const int64_t f = i / nch;
    const int s = (int)(i - f * nch);
    FrameDelayDev d;
    compute_frame_delay(a.first_frame + f, a.dly, d, nullptr, nullptr);
    float2 p, step;
    {
#pragma clang fp contract(off)
        if (kind == CHAIN_XCF) {
            const double two_pi_df = -2.0 * FXC_PI * a.df;
            const double step1 = two_pi_df * d.frac1;
            const double step2 = two_pi_df * d.frac2;
            p = polar1_f32((float)(step1 * (double)a.a1s - step2 * (double)a.a2s));
            step = polar1_f32((float)(step1 - step2));
        } else {
            const double total = s ? d.frac2 : d.frac1;
            const double phase_step = -1.0 * 2.0 * FXC_PI * a.df * total;
            p = polar1_f32((float)(phase_step * 0.0));
            step = polar1_f32((float)phase_step);
        }
    }
    // row of a.chain_ld (even, >= len + (len & 1)) float2: 16-byte aligned
    float4 *o = reinterpret_cast<float4 *>(out + i * (int64_t)a.chain_ld);
    for (int k = 0; k < len; k += 2) {
        const float2 q = c_mul_rust(p, step);
        o[k >> 1] = make_float4(p.x, p.y, q.x, q.y);
        p = c_mul_rust(q, step);
    }
}

hipError_t launch_phase_chain(const SectorArgs &a, int kind, float2 *out, hipStream_t s) {
    const int64_t n = a.nf * (kind == CHAIN_XCF ? 1 : 2);
    if (n <= 0) return hipSuccess;
    const int len = kind == CHAIN_XCF ? a.ov : a.M + 1;
    if (a.chain_ld < chain_ld_for(len) || (a.chain_ld & 1)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_phase_chain, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, a, kind,
                       out);
    return hipGetLastError();
}

}  // namespace fxc
