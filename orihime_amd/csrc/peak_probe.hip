// Issue-rate probes for the roofline of the soft Smith-Waterman sweeps.
//
// The sweeps are bound by vector-ALU issue: FP32 add/max/fma, DPP lane
// moves and the transcendental v_exp_f32 / v_log_f32.  Published MI355X
// figures cover the packed-FP32 and matrix peaks only, so these kernels
// measure the per-instruction rates on the device: each lane runs 8
// independent chains of one instruction kind (no dependency stalls at
// 8 waves per SIMD), and the host divides lane-operations by the elapsed
// time of the launch.
#include <hip/hip_runtime.h>
#include <cstdint>

#include "ohm_probe.h"

namespace {

constexpr int kChains = 8;
constexpr int kThreads = 256;

template <int KIND>
__global__ __launch_bounds__(kThreads) void probe_kernel(float* __restrict__ sink, int iters,
                                                         float a, float b) {
    const int gid = blockIdx.x * kThreads + threadIdx.x;
    float x[kChains], y[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) {
        x[c] = 0.25f + 1e-3f * (float)((gid + c) & 63);
        y[c] = x[c];
    }
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) {
            if constexpr (KIND == OHM_PROBE_EXP) {
                x[c] = __builtin_amdgcn_exp2f(-x[c]);
            } else if constexpr (KIND == OHM_PROBE_LOG) {
                x[c] = __builtin_amdgcn_logf(__builtin_fabsf(x[c]));  // |x| is a source modifier
            } else if constexpr (KIND == OHM_PROBE_FMA) {
                x[c] = __builtin_fmaf(x[c], a, b);
            } else if constexpr (KIND == OHM_PROBE_DPP) {
                // wave shift right by one lane (the sweeps' neighbour move)
                x[c] = __builtin_bit_cast(
                    float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, x[c]), __builtin_bit_cast(int, x[c]), 0x138,
                                                       0xf, 0xf, false));
            } else {  // OHM_PROBE_MIX: one exp and four independent FMAs
                x[c] = __builtin_amdgcn_exp2f(-x[c]);
                y[c] = __builtin_fmaf(y[c], a, b);
                y[c] = __builtin_fmaf(y[c], a, b);
                y[c] = __builtin_fmaf(y[c], a, b);
                y[c] = __builtin_fmaf(y[c], a, b);
            }
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < kChains; ++c) s += x[c] + y[c];
    sink[gid] = s;
}

}  // namespace

extern "C" int ohm_probe_ops_per_thread(int kind, int iters, int64_t* ops) {
    if (!ops || iters < 0) return 1;
    const int64_t per = (int64_t)iters * kChains;
    switch (kind) {
        case OHM_PROBE_EXP:
        case OHM_PROBE_LOG:
        case OHM_PROBE_FMA:
        case OHM_PROBE_DPP: *ops = per; return 0;
        case OHM_PROBE_MIX: *ops = 5 * per; return 0;
        default: return 1;
    }
}

extern "C" int ohm_probe_launch(int kind, int blocks, int iters, float* sink, void* stream) {
    if (blocks < 1 || iters < 0 || !sink) return 1;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const dim3 g(blocks), t(kThreads);
    switch (kind) {
        case OHM_PROBE_EXP: hipLaunchKernelGGL(probe_kernel<OHM_PROBE_EXP>, g, t, 0, st, sink, iters, 1.0f, 0.0f); break;
        case OHM_PROBE_LOG: hipLaunchKernelGGL(probe_kernel<OHM_PROBE_LOG>, g, t, 0, st, sink, iters, 1.0f, 0.0f); break;
        case OHM_PROBE_FMA: hipLaunchKernelGGL(probe_kernel<OHM_PROBE_FMA>, g, t, 0, st, sink, iters, 0.999f, 1e-3f); break;
        case OHM_PROBE_DPP: hipLaunchKernelGGL(probe_kernel<OHM_PROBE_DPP>, g, t, 0, st, sink, iters, 1.0f, 0.0f); break;
        case OHM_PROBE_MIX: hipLaunchKernelGGL(probe_kernel<OHM_PROBE_MIX>, g, t, 0, st, sink, iters, 0.999f, 1e-3f); break;
        default: return 1;
    }
    return hipGetLastError() == hipSuccess ? 0 : 2;
}
