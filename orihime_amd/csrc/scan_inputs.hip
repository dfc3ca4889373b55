// Input scan of the drop-in validation (include/ohm_sw.h, ohm_scan_inputs).
//
// The reference validates a call with several torch reductions and host
// copies: the numerical domain and the mask marker (orihime/operator.py
// :249-386: no +inf, no NaN, every finite score within `bound` of 0 unless it
// is the legacy marker), the mask sentinel's per-pair maximum (:334-386), and
// the lengths range (src/sw_affine/torch_hip.cpp:69-99).  Here one pass over
// the padded scores gathers all of it into four counts and the per-pair
// maxima, so the host reads back one small block once per call.
//
// Three launches, no host synchronisation, no memset (graph-capturable):
// scan_init clears the counts and seeds the maxima, scan_kernel reduces each
// block of a pair's cells in registers and merges with integer atomics
// (counts: adds; maxima: unsigned max over an order-preserving key of the
// float bits, so the result does not depend on the order), scan_finish turns
// the keys back into floats.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "ohm_sw.h"

namespace ohm {
void set_error(const std::string& msg);
}

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanPerThread = 16;  // floats per thread per block (one coalesced row of the block per step)

// Unsigned key with the order of the float values (NaN excluded).
__device__ __forceinline__ uint32_t order_key(float x) {
    const uint32_t u = __builtin_bit_cast(uint32_t, x);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_value(uint32_t k) {
    const uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __builtin_bit_cast(float, u);
}

__device__ __forceinline__ int wave_add(int v) {
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ float wave_fmax(float v) {
    for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

// counts: [0] +inf cells, [1] NaN or out-of-domain finite cells, [2] -inf
// cells, [3] lengths entries outside [0, L1] x [0, L2].
__global__ __launch_bounds__(256) void scan_init(int32_t* counts, uint32_t* keys, int B) {
    const int t = threadIdx.x;
    if (t < 4) counts[t] = 0;
    const uint32_t ninf = order_key(-__builtin_inff());
    for (int b = t; b < B; b += blockDim.x) keys[b] = ninf;
}

// One-dimensional grid of B * nblk blocks: block x scans
// kScanThreads * kScanPerThread consecutive cells of pair x / nblk (whole
// pair boxes of the padded layout: the cells past a pair's lengths are
// scanned too, as the reference does).
__global__ __launch_bounds__(kScanThreads) void scan_kernel(const float* __restrict__ scores,
                                                            const int32_t* __restrict__ lengths,
                                                            int64_t cells, int nblk, int L1, int L2,
                                                            float bound, float marker,
                                                            int32_t* counts, uint32_t* keys) {
    const int b = blockIdx.x / nblk, blk = blockIdx.x % nblk;
    const int t = threadIdx.x;
    if (blk == 0 && t == 0 && lengths) {
        const int l1 = lengths[2 * b], l2 = lengths[2 * b + 1];
        const int bad = (l1 < 0 || l1 > L1) + (l2 < 0 || l2 > L2);
        if (bad) atomicAdd(&counts[3], bad);
    }
    const float* base = scores + (int64_t)b * cells;
    const int64_t start = (int64_t)blk * (kScanThreads * kScanPerThread);
    int pinf = 0, bad = 0, ninf = 0;
    float mx = -__builtin_inff();
#pragma unroll
    for (int q = 0; q < kScanPerThread; ++q) {
        const int64_t i = start + (int64_t)q * kScanThreads + t;  // coalesced per q
        if (i < cells) {
            // classified on the bits: the library is built without NaN
            // semantics for float compares (-fno-honor-nans)
            const float x = base[i];
            const uint32_t u = __builtin_bit_cast(uint32_t, x), m = u & 0x7fffffffu;
            const bool is_nan = m > 0x7f800000u;
            const bool is_pinf = u == 0x7f800000u;
            const bool is_ninf = u == 0xff800000u;
            pinf += is_pinf;
            ninf += is_ninf;
            bad += is_nan || (m < 0x7f800000u && x != marker && fabsf(x) > bound);
            if (!is_nan) mx = fmaxf(mx, x);
        }
    }
    pinf = wave_add(pinf);
    bad = wave_add(bad);
    ninf = wave_add(ninf);
    mx = wave_fmax(mx);
    if ((t & 63) == 0) {
        if (pinf) atomicAdd(&counts[0], pinf);
        if (bad) atomicAdd(&counts[1], bad);
        if (ninf) atomicAdd(&counts[2], ninf);
        if (mx > -__builtin_inff()) atomicMax(&keys[b], order_key(mx));
    }
}

__global__ __launch_bounds__(256) void scan_finish(uint32_t* keys, int B) {
    for (int b = threadIdx.x; b < B; b += blockDim.x) {
        const float v = key_value(keys[b]);
        reinterpret_cast<float*>(keys)[b] = v;
    }
}

ohm_status fail(ohm_status st, const std::string& msg) {
    ohm::set_error(msg);
    return st;
}

}  // namespace

extern "C" ohm_status ohm_scan_inputs(const float* scores, const int32_t* lengths, int32_t B,
                                      int32_t L1, int32_t L2, float bound, float marker,
                                      int32_t* counts, float* pair_max, void* stream) {
    if (B < 0 || L1 < 0 || L2 < 0) return fail(OHM_EINVAL, "negative batch size or length");
    if (!counts) return fail(OHM_EINVAL, "counts is null");
    if (B > 0 && !pair_max) return fail(OHM_EINVAL, "pair_max is null");
    const int64_t cells = (int64_t)L1 * L2;
    if (B > 0 && cells > 0 && !scores) return fail(OHM_EINVAL, "scores is null");
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint32_t* keys = reinterpret_cast<uint32_t*>(pair_max);
    hipLaunchKernelGGL(scan_init, dim3(1), dim3(256), 0, s, counts, keys, B);
    if (B > 0) {
        const int64_t per_block = (int64_t)kScanThreads * kScanPerThread;
        const int64_t nblk = cells > 0 ? (cells + per_block - 1) / per_block : 1;
        if (nblk * B > 0x7fffffff) return fail(OHM_EINVAL, "batch too large for one scan launch");
        hipLaunchKernelGGL(scan_kernel, dim3((unsigned)(nblk * B)), dim3(kScanThreads), 0, s,
                           scores, lengths, cells, (int)nblk, L1, L2, bound, marker, counts, keys);
        hipLaunchKernelGGL(scan_finish, dim3(1), dim3(256), 0, s, keys, B);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(OHM_ELAUNCH, std::string("scan kernels: ") + hipGetErrorString(e));
    return OHM_OK;
}
