// Soft Smith-Waterman (linear and affine gap) for AMD Instinct MI355X (gfx950).
//
// Work decomposition
//   A pair's L1 x L2 score matrix is cut into horizontal strips of 64 rows.
//   One wavefront sweeps one strip: lane r owns row 64*s + r + 1 and, at step
//   t, computes column j = t - r + 1, so a step is one anti-diagonal of the
//   strip and every lane stays busy after a 63-step ramp.  Neighbour values
//   move between rows with DPP wave_shr:1 / wave_shl:1 (lane r <- r-1 / r+1),
//   the left neighbour is the lane's own previous step.  Lane 0 receives the
//   row above from the previous strip's last row, which that strip's lane 63
//   writes to a per-pair boundary block in global memory; a per-strip
//   progress counter in LDS (release/acquire at workgroup scope) tells the
//   consumer how many boundary columns are ready.
//   A workgroup of W waves owns one pair at a time; its waves take the pair's
//   strips round-robin, so strips of a pair run concurrently with a lag of
//   about one chunk per strip.  Workgroups are persistent: they claim pairs
//   from a device counter in largest-first order (the plan kernel sorts the
//   pairs by cost), so a ragged batch runs in one launch with no padding.
//
// Memory traffic
//   Scores are read per lane in 16-column chunks that are skewed by the lane
//   index (lane r's chunk k covers columns 16k - r .. 16k - r + 15), so each
//   step uses a compile-time register index and each lane issues 64-byte
//   vector loads.  The forward pass stores the DP table of a strip in a skewed
//   layout [step][plane][lane]; each store and each backward load is one
//   coalesced 256-byte access per plane.
//
// Derivatives
//   The backward pass is the pull form of the reference's push adjoint
//   (src/sw_affine/kernels_cpu.cpp:201-356): a cell reads the adjoints of its
//   successors instead of scattering into its predecessors, so no atomics are
//   needed and every reduction has a fixed order (bitwise repeatable).
//   Forward-over-reverse passes (HVP, parameter sensitivities, map
//   cotangents) propagate a tangent through the forward sweep and the
//   adjoint's tangent through the reverse sweep.
//
// Numerics: FP32 max-shifted log-sum-exp in log2 units (see sw_common.h).

#include <hip/hip_runtime.h>

#include <atomic>
#include <stdint.h>
#include <stdio.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <string>

#include "ohm_sw.h"
#include "sw_common.h"

using namespace ohm;

#define DEV static __device__ __forceinline__

// --------------------------------------------------------------------------
// Device helpers
// --------------------------------------------------------------------------

// Work-item index within the workgroup (read from the hardware register; the
// threadIdx path goes through an out-of-line library call).
DEV int tid_x() { return (int)__builtin_amdgcn_workitem_id_x(); }
DEV float ex2(float x) { return __builtin_amdgcn_exp2f(x); }
DEV float lg2(float x) { return __builtin_amdgcn_logf(x); }
DEV float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }
DEV float med3(float a, float b, float c) { return __builtin_amdgcn_fmed3f(a, b, c); }

// Lane l receives lane l-1's value; lane 0 receives `edge`  (DPP wave_shr:1).
DEV float from_prev_lane(float v, float edge) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, edge),
                                           __builtin_bit_cast(int, v), 0x138, 0xf, 0xf, false));
}
// Lane l receives lane l+1's value; lane 63 receives `edge`  (DPP wave_shl:1).
DEV float from_next_lane(float v, float edge) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, edge),
                                           __builtin_bit_cast(int, v), 0x130, 0xf, 0xf, false));
}
// Boundary queues: the 16 boundary entries of a chunk sit in lanes 0..15
// (forward, entry u in lane u) or 48..63 (reverse, entry u in lane 63 - u);
// one rotation per step brings entry u to lane 0 / lane 63 at step u, where
// from_prev_lane / from_next_lane pick it up as their edge value.
DEV float rotate_down(float q) {  // lane l <- lane l+1 (lane 63 <- 0)
    return __builtin_bit_cast(
        float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, q), 0x130, 0xf, 0xf, true));
}
DEV float rotate_up(float q) {  // lane l <- lane l-1 (lane 0 <- 0)
    return __builtin_bit_cast(
        float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, q), 0x138, 0xf, 0xf, true));
}
// True when a > b in any active lane.  The compare writes a scalar lane
// mask directly, so the common (false) case costs one vector compare and a
// scalar branch.
DEV bool any_greater(float a, float b) {
    uint64_t m;
    asm volatile("v_cmp_gt_f32_e64 %0, %1, %2" : "=s"(m) : "v"(a), "v"(b));
    return m != 0;
}
// One step of a queue: returns the neighbour shift of v with the queue's
// current head as edge value and advances the queue.  The rotation is issued
// first so the shift can overwrite the old queue register in place.
DEV float shift_in_prev(float v, float& q) {
    const float nq = rotate_down(q);
    const float x = from_prev_lane(v, q);
    q = nq;
    return x;
}
DEV float shift_in_next(float v, float& q) {
    const float nq = rotate_up(q);
    const float x = from_next_lane(v, q);
    q = nq;
    return x;
}
// The same moves for 32-bit integers (the scaled sweeps' exponents).
DEV int shift_in_prev_i(int v, int& q) {
    const int nq = __builtin_amdgcn_mov_dpp(q, 0x130, 0xf, 0xf, true);
    const int x = __builtin_amdgcn_update_dpp(q, v, 0x138, 0xf, 0xf, false);
    q = nq;
    return x;
}
DEV int shift_in_next_i(int v, int& q) {
    const int nq = __builtin_amdgcn_mov_dpp(q, 0x138, 0xf, 0xf, true);
    const int x = __builtin_amdgcn_update_dpp(q, v, 0x130, 0xf, 0xf, false);
    q = nq;
    return x;
}
DEV bool any_greater_i(int a, int b) {
    uint64_t m;
    asm volatile("v_cmp_gt_i32_e64 %0, %1, %2" : "=s"(m) : "v"(a), "v"(b));
    return m != 0;
}
// x * 2^e and the binary exponent of x (x = m * 2^e, m in [0.5, 1)), exact.
DEV float ldx(float x, int e) { return __builtin_amdgcn_ldexpf(x, e); }
DEV int fexp(float x) { return __builtin_amdgcn_frexp_expf(x); }
// Finite test on the bit pattern (the device code is built without NaN
// semantics, so isfinite() may be folded away).
DEV bool finite_bits(float x) { return (__builtin_bit_cast(uint32_t, x) & 0x7f800000u) != 0x7f800000u; }

// Global accesses as (wave-uniform base pointer) + (32-bit lane offset): the
// compiler emits the scalar-base form, so per-step address updates are scalar
// adds and cost no vector ALU.
DEV float gload(const float* base, uint32_t byte_off) {
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + byte_off);
}
DEV void gstore(float* base, uint32_t byte_off, float v) {
    *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + byte_off) = v;
}

DEV float lane_value(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}

DEV float wave_sum(float v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}
DEV float wave_max(float v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = fmaxf(v, __shfl_xor(v, m, 64));
    return v;
}

// log2(2^a + 2^b)
DEV float lse2(float a, float b) {
    float hi = fmaxf(a, b), lo = fminf(a, b);
    return hi + lg2(1.0f + ex2(lo - hi));
}
// log2(2^a + 2^b + 2^c): the largest term contributes exactly 1.
DEV float lse3(float a, float b, float c) {
    float hi = fmaxf(fmaxf(a, b), c), lo = fminf(fminf(a, b), c), md = med3(a, b, c);
    return hi + lg2(1.0f + ex2(md - hi) + ex2(lo - hi));
}
// log2(1 + 2^x) and its derivative 2^x / (1 + 2^x)
DEV float softplus2(float x) { return fmaxf(x, 0.0f) + lg2(1.0f + ex2(-fabsf(x))); }
DEV float sigmoid2(float x) { return rcpf(1.0f + ex2(-x)); }
// sigmoid2(x) and 1 - sigmoid2(x), both without cancellation.
DEV void sigmoid2_pair(float x, float& w, float& wc) {
    float e = ex2(-fabsf(x));
    float r = rcpf(1.0f + e);
    w = x >= 0.0f ? r : e * r;
    wc = x >= 0.0f ? e * r : r;
}
DEV void softplus_sigmoid2(float x, float& sp, float& sg) {
    float e = ex2(-fabsf(x));
    float r = rcpf(1.0f + e);
    sp = fmaxf(x, 0.0f) + lg2(1.0f + e);
    sg = x >= 0.0f ? r : e * r;
}
// Two normalised weights 2^la, 2^lb with 2^la + 2^lb = 1 up to rounding: the
// smaller one is evaluated directly, the larger as its complement, so both
// keep full relative precision.  Log-weights are clamped at 0 so that an
// absent successor (adjoint 0) can never produce inf * 0.
DEV void pair_weights(float la, float lb, float& wa, float& wb) {
    float lo = fminf(fminf(la, lb), 0.0f);
    float wl = ex2(lo), wh = 1.0f - wl;
    bool a_lo = la <= lb;
    wa = a_lo ? wl : wh;
    wb = a_lo ? wh : wl;
}
// y = log2(2^a + 2^b + 2^c) and the three weights, normalised by their own
// sum.  Taking 2^(x - y) with a rounded y instead would lose a relative
// ~ulp(y) of adjoint flow at every node, which compounds along a path.
DEV void lse3_norm(float a, float b, float c, float& y, float& wa, float& wb, float& wc) {
    const float hi = fmaxf(fmaxf(a, b), c);
    const float ea = ex2(a - hi), eb = ex2(b - hi), ec = ex2(c - hi);
    const float sum = ea + eb + ec, r = rcpf(sum);
    y = hi + lg2(sum);
    wa = ea * r; wb = eb * r; wc = ec * r;
}

// 16 consecutive floats row[c0 .. c0+15]; 0 outside [0, len) or when !ok.
// Fast path: four unaligned 16-byte loads (dword alignment suffices).
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
DEV void load16(float (&dst)[16], const float* row, int c0, int len, bool ok) {
    if (ok && c0 >= 0 && c0 + 16 <= len) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            f4u v = *reinterpret_cast<const f4u*>(row + c0 + 4 * q);
            dst[4 * q + 0] = v.x; dst[4 * q + 1] = v.y; dst[4 * q + 2] = v.z; dst[4 * q + 3] = v.w;
        }
    } else {
        // one window base, element u at an immediate offset (a per-element
        // 64-bit row index would hold 16 address pairs live)
        const float* p = row + c0;
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            int c = c0 + u;
            dst[u] = (ok && c >= 0 && c < len) ? p[u] : 0.0f;
        }
    }
}
DEV void store16(float* row, int c0, int len, bool ok, const float (&src)[16]) {
    if (ok && c0 >= 0 && c0 + 16 <= len) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            f4u v;
            v.x = src[4 * q + 0]; v.y = src[4 * q + 1]; v.z = src[4 * q + 2]; v.w = src[4 * q + 3];
            *reinterpret_cast<f4u*>(row + c0 + 4 * q) = v;
        }
    } else {
        float* p = row + c0;
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            int c = c0 + u;
            if (ok && c >= 0 && c < len) p[u] = src[u];
        }
    }
}

// 8-float versions (two unaligned 16-byte accesses on the fast path).
DEV void load8(float (&dst)[8], const float* row, int c0, int len, bool ok) {
    if (ok && c0 >= 0 && c0 + 8 <= len) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            f4u v = *reinterpret_cast<const f4u*>(row + c0 + 4 * q);
            dst[4 * q + 0] = v.x; dst[4 * q + 1] = v.y; dst[4 * q + 2] = v.z; dst[4 * q + 3] = v.w;
        }
    } else {
        const float* p = row + c0;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            int c = c0 + u;
            dst[u] = (ok && c >= 0 && c < len) ? p[u] : 0.0f;
        }
    }
}

// K consecutive floats of a row (K = 4, 8 or 16: the checkpoint block width)
template <int K>
DEV void load_run(float (&dst)[K], const float* row, int c0, int len, bool ok) {
    if constexpr (K == 8) {
        load8(dst, row, c0, len, ok);
    } else if constexpr (K == 16) {
        load16(dst, row, c0, len, ok);
    } else {
        const float* p = row + c0;
#pragma unroll
        for (int u = 0; u < K; ++u) {
            const int c = c0 + u;
            dst[u] = (ok && c >= 0 && c < len) ? p[u] : 0.0f;
        }
    }
}
DEV void store8(float* row, int c0, int len, bool ok, const float (&src)[8]) {
    if (ok && c0 >= 0 && c0 + 8 <= len) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            f4u v;
            v.x = src[4 * q + 0]; v.y = src[4 * q + 1]; v.z = src[4 * q + 2]; v.w = src[4 * q + 3];
            *reinterpret_cast<f4u*>(row + c0 + 4 * q) = v;
        }
    } else {
        float* p = row + c0;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            int c = c0 + u;
            if (ok && c >= 0 && c < len) p[u] = src[u];
        }
    }
}

// Strip hand-off.  Progress counters live in LDS; boundary data in global
// memory written by the producer wave before the release.
DEV void wait_ready(int* flag, int need) {
    while (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < need)
        __builtin_amdgcn_s_sleep(1);
}
DEV void publish(int* flag, int value, int lane) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if (lane == 0) __hip_atomic_store(flag, value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Hand-off between waves of different workgroups (grid scheduling).
//
// Boundary entries are NaN-boxed: before a grid launch every word of the
// pairs' boundary blocks is filled with kGridEmpty, a signalling-NaN pattern
// that arithmetic never produces (invalid operations give the default quiet
// NaN and NaN operands propagate quieted; scores are NaN-free).  A producer
// stores its entries with agent-scope atomic stores (global_store sc1:
// coherent per access across CUs and XCDs) and publishes nothing else; a
// consumer takes a word as ready once an agent-scope load returns something
// other than kGridEmpty (32-bit stores are single-copy atomic).  No progress
// words, no fences and no store-completion waits: consumers prefetch the
// next chunk's entries while they compute the current chunk and poll again
// only for words still empty.  A poll is bounded: past kGridPollCap rounds,
// or once another wave has given up, it raises claims[kClaimGridStall] and
// returns (the outputs are then NaN; the ticket order makes every producer a
// running wave, so no wait comes near the bound).
constexpr uint32_t kGridEmpty = 0x7fbadbadu;
constexpr int kGridPollCap = 1 << 20;
DEV float cload(const float* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
DEV void cstore(float* p, float v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
DEV bool grid_empty(float v) { return __builtin_bit_cast(uint32_t, v) == kGridEmpty; }

// Wait until the N words w of every lane with `need` are filled; fetch(w)
// reloads a lane's words.
template <int N, typename Fetch>
DEV void grid_ready(float (&w)[N], bool need, Fetch&& fetch, int* stall) {
    for (int n = 0;; ++n) {
        bool empty = false;
#pragma unroll
        for (int i = 0; i < N; ++i) empty |= grid_empty(w[i]);
        if (__builtin_amdgcn_ballot_w64(need && empty) == 0) break;
        if (n >= kGridPollCap ||
            ((n & 255) == 255 && __hip_atomic_load(stall, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
            __hip_atomic_store(stall, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
        if (need && empty) fetch(w);
    }
}

// Boundary entry store: plain (one workgroup, ordered by the LDS progress
// words) or agent-scope atomic (grid scheduling).
template <bool GRID>
DEV void hand_store(float* p, float v) {
    if constexpr (GRID) cstore(p, v);
    else *p = v;
}

// Kahan add of a chunk partial into a running total.
DEV void kadd(float& sum, float& comp, float v) {
    float y = v - comp;
    float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

// Boundary columns consumed at chunk k: entry index (16k + u) for u = 0..15,
// held by lane u.  Ready when the producer published >= min(L2, 16k + 16).
DEV int chunk_need(int k, int L2) { return min(L2, kChunk * k + kChunk); }
// Entries a producer strip has completed after finishing chunk k: its lane 63
// (lane 0 in reverse) lags lane 0 by 63 steps.
DEV int chunk_done(int k, int L2) { return max(0, min(L2, kChunk * k + kChunk - 63)); }

// --------------------------------------------------------------------------
// Plan: per-pair descriptors, exact table/boundary offsets, largest-first
// order, claim counters.  One block of 1024 threads.
// --------------------------------------------------------------------------

constexpr int kPlanThreads = 1024;
constexpr int kSortMax = 2048;

__global__ __launch_bounds__(kPlanThreads) void sw_plan_kernel(
    const int32_t* __restrict__ lengths, const int64_t* __restrict__ offsets, int B, int L1max,
    int L2max, int planes, int tstride, int64_t cap, PairDesc* __restrict__ desc,
    int32_t* __restrict__ order, int* __restrict__ claims, int* __restrict__ grid,
    int grid_ints) {
    __shared__ int64_t s_t[kPlanThreads];
    __shared__ int64_t s_carry[1];
    __shared__ float s_key[kSortMax];
    __shared__ int s_idx[kSortMax];
    __shared__ int s_differs[kPlanThreads / kWave];  // per wave: a key differs from pair 0's
    const int tid = tid_x();
    if (tid < 64) claims[tid] = 0;
    if (grid)
        for (int q = tid; q < grid_ints; q += kPlanThreads) grid[q] = 0;
    if (tid == 0) s_carry[0] = 0;
    __syncthreads();
    const bool sort = B <= kSortMax;
    for (int base = 0; base < B; base += kPlanThreads) {
        const int b = base + tid;
        int L1 = 0, L2 = 0, ns = 0;
        int64_t ts = 0, bs = 0;
        if (b < B) {
            L1 = lengths ? lengths[2 * b] : L1max;
            L2 = lengths ? lengths[2 * b + 1] : L2max;
            // Lengths outside [0, max] never address outside the scores or the
            // LDS progress counters (sized for ceil(L1max / 64) strips).
            L1 = min(max(L1, 0), L1max);
            L2 = min(max(L2, 0), L2max);
            ns = (L1 > 0 && L2 > 0) ? (L1 + kStrip - 1) / kStrip : 0;
            ts = (int64_t)ns * ((L2 + 63 + tstride - 1) / tstride) * kWave * planes;
            bs = (int64_t)ns * L2 * kBndStride;
        }
        // each pair's region starts on a 256-byte boundary, so the 256-byte
        // table rows of its strips fill whole cache lines
        const int64_t tot = (ts + bs + 63) / 64 * 64;
        s_t[tid] = tot;
        __syncthreads();
        for (int d = 1; d < kPlanThreads; d <<= 1) {
            int64_t at = tid >= d ? s_t[tid - d] : 0;
            __syncthreads();
            s_t[tid] += at;
            __syncthreads();
        }
        if (b < B) {
            PairDesc pd;
            pd.t_off = s_carry[0] + s_t[tid] - tot;
            pd.b_off = pd.t_off + ts;
            pd.s_off = offsets ? offsets[b] : (int64_t)b * L1max * L2max;
            pd.ld = offsets ? L2 : L2max;
            pd.L1 = L1;
            pd.L2 = L2;
            pd.ns = ns;
            pd.lost = 0;
            if (pd.b_off + bs > cap) {
                // Workspace sized for fewer cells than the lengths ask for (a
                // packed-layout call): the pair runs as an empty pair and
                // lost_pairs_kernel, launched at the end of the call, turns its
                // outputs into NaN.  Every writer stores the same word.
                pd.ns = 0;
                pd.L1 = 0;
                pd.L2 = 0;
                pd.lost = 1;
                claims[kClaimWsShort] = 1;
            }
            desc[b] = pd;
            if (sort) {
                s_key[b] = (float)ns * (float)(L2 + 63);
                s_idx[b] = b;
            } else {
                order[b] = b;
            }
        }
        __syncthreads();
        if (tid == kPlanThreads - 1) s_carry[0] += s_t[tid];
        __syncthreads();
    }
    if (!sort) return;
    {
        bool differs = false;
        for (int q = tid; q < B; q += kPlanThreads) differs |= s_key[q] != s_key[0];
        const uint64_t m = __builtin_amdgcn_ballot_w64(differs);
        if ((tid & (kWave - 1)) == 0) s_differs[tid / kWave] = m != 0;
        __syncthreads();
        bool any = false;
        for (int w = 0; w < kPlanThreads / kWave; ++w) any |= s_differs[w] != 0;
        if (!any) {
            // every pair costs the same: the order (ties by index) is the identity
            for (int q = tid; q < B; q += kPlanThreads) order[q] = q;
            return;
        }
    }
    int n = 1;
    while (n < B) n <<= 1;
    for (int q = B + tid; q < n; q += kPlanThreads) {
        s_key[q] = -1.0f;
        s_idx[q] = q;
    }
    __syncthreads();
    // Bitonic sort, descending by cost, ties by index.  Stages whose partners
    // are 64 or more positions apart exchange through LDS (a barrier each);
    // the rest of a level runs in registers, partners exchanged by lane
    // shuffles (position q = tid + 1024 m lives in lane tid % 64 of wave
    // tid / 64, its partner q ^ j, j < 64, in the same wave).  Every stage
    // swaps partner positions, so the result is a permutation.
    constexpr int kPer = kSortMax / kPlanThreads;
    for (int k = 2; k <= n; k <<= 1) {
        int j = k >> 1;
        for (; j >= kWave; j >>= 1) {
            for (int q = tid; q < n; q += kPlanThreads) {
                int p = q ^ j;
                if (p > q) {
                    bool desc_block = (q & k) == 0;
                    float kq = s_key[q], kp = s_key[p];
                    int iq = s_idx[q], ip = s_idx[p];
                    bool q_first = kq > kp || (kq == kp && iq < ip);
                    if (q_first != desc_block) {
                        s_key[q] = kp; s_key[p] = kq;
                        s_idx[q] = ip; s_idx[p] = iq;
                    }
                }
            }
            __syncthreads();
        }
        float rk[kPer];
        int ri[kPer];
#pragma unroll
        for (int m = 0; m < kPer; ++m) {
            const int q = tid + kPlanThreads * m;
            rk[m] = q < n ? s_key[q] : -1.0f;
            ri[m] = q < n ? s_idx[q] : q;
        }
        const int wbase = tid & ~(kWave - 1);  // the wave's first position (m = 0)
        for (; j > 0; j >>= 1) {
#pragma unroll
            for (int m = 0; m < kPer; ++m) {
                // a wave whose positions are all padding skips them (wave-uniform)
                if (wbase + kPlanThreads * m >= n) continue;
                const int q = tid + kPlanThreads * m;
                const float kp = __shfl_xor(rk[m], j, kWave);
                const int ip = __shfl_xor(ri[m], j, kWave);
                const bool desc_block = ((q & ~j) & k) == 0;  // block of the pair's lower position
                const bool mine_first = rk[m] > kp || (rk[m] == kp && ri[m] < ip);
                const bool lower = (q & j) == 0;
                // the lower position takes the first element of a descending
                // block and the second of an ascending one
                const bool keep = lower ? (mine_first == desc_block) : (mine_first != desc_block);
                if (!keep) {
                    rk[m] = kp;
                    ri[m] = ip;
                }
            }
        }
#pragma unroll
        for (int m = 0; m < kPer; ++m) {
            const int q = tid + kPlanThreads * m;
            if (q < n) {
                s_key[q] = rk[m];
                s_idx[q] = ri[m];
            }
        }
        __syncthreads();
    }
    for (int q = tid; q < B; q += kPlanThreads) order[q] = s_idx[q];
}

// Zero the padded-layout cells of pair b outside its lengths, with the
// `nthr` threads of the caller (a wave or a workgroup): the tail of each of
// the rows [0, L1) row by row, threads across its columns, then the
// contiguous rows [L1, L1max).  (No per-element index division: a padded
// pair can have millions of such cells.)
DEV void zero_padding(float* out, const PairDesc& d, int L1max, int L2max, int tid, int nthr) {
    if (!out) return;
    float* base = out + d.s_off;
    const int L1 = d.L1, L2 = d.L2;
    const int w = L2max - L2;
    if (w > 0)
        for (int rr = 0; rr < L1; ++rr) {
            float* row = base + (int64_t)rr * L2max + L2;
            for (int c = tid; c < w; c += nthr) row[c] = 0.0f;
        }
    float* p = base + (int64_t)L1 * L2max;
    const int64_t n = (int64_t)(L1max - L1) * L2max;
    for (int64_t q = tid; q < n; q += nthr) p[q] = 0.0f;
}

// Outputs of one call, for lost_pairs_kernel.
struct CallOut {
    float* value;
    float* g[3];
    float* contract;  // [B * nc]
    int nc;
    float* map;
    float* map_t;
};

// End of a packed-layout call whose workspace was sized for fewer cells than
// the lengths ask for: the plan kernel ran the pairs that did not fit as empty
// pairs and raised claims[kClaimWsShort]; this one-block launch overwrites
// their value, gradients, contractions and every map / map-tangent cell with
// NaN, so the failure cannot pass for an empty alignment.  It returns at once
// when the word is clear.  Lengths are clamped to [0, max] as the plan kernel
// clamps them, so the cells written are the ones the pair owns.
__global__ __launch_bounds__(256) void lost_pairs_kernel(const int* __restrict__ claims,
                                                         const PairDesc* __restrict__ desc,
                                                         const int32_t* __restrict__ lengths,
                                                         const int64_t* __restrict__ offsets,
                                                         int B, int L1max, int L2max, CallOut o) {
    if (__builtin_amdgcn_readfirstlane(claims[kClaimWsShort]) == 0) return;
    const float nan = __builtin_nanf("");
    const int tid = tid_x();
    for (int b = tid; b < B; b += 256) {
        if (!desc[b].lost) continue;
        if (o.value) o.value[b] = nan;
        for (int q = 0; q < 3; ++q)
            if (o.g[q]) o.g[q][b] = nan;
        if (o.contract)
            for (int q = 0; q < o.nc; ++q) o.contract[(int64_t)o.nc * b + q] = nan;
    }
    if (!o.map && !o.map_t) return;
    for (int b = 0; b < B; ++b) {
        if (!desc[b].lost) continue;
        const int64_t n = (int64_t)min(max(lengths[2 * b], 0), L1max) *
                          min(max(lengths[2 * b + 1], 0), L2max);
        for (int64_t q = tid; q < n; q += 256) {
            if (o.map) o.map[offsets[b] + q] = nan;
            if (o.map_t) o.map_t[offsets[b] + q] = nan;
        }
    }
}

// --------------------------------------------------------------------------
// Kernel arguments (passed by value: graph-capturable)
// --------------------------------------------------------------------------

struct SwArgs {
    const float* scores;     // pair scores
    const float* dir;        // score-space tangent direction (tangent passes) or null
    const PairDesc* desc;
    const int32_t* order;
    int* claims;             // claim counters, zeroed by the plan kernel
    float* table;            // DP tables (skewed layout)
    float* bnd;              // strip boundary blocks
    float* stil;             // per pair: partition in log2 units
    float* tgt;              // per pair: 4 floats of tangent-pass scalars
    float* value;            // [B] partition (natural units)
    float* map;              // map output (nullable)
    float* map_t;            // map tangent output (nullable)
    float* g0;               // [B] gradient outputs (nullable)
    float* g1;
    float* g2;
    float* contract;         // [B * nparams] contracted tangents (nullable)
    int B, L1max, L2max;
    int padded;              // 1: padded layout (zero-fill outside lengths)
    int max_strips;          // LDS progress counters needed
    float kappa;             // log2(e) / T
    float go2, ge2;          // gap_open, gap_ext (affine) or gap (linear, go2) in log2 units
    float go, ge, temp;      // natural units
    float d_go, d_ge, d_temp;  // scalar tangent directions
    float lgo, lge;          // 2^go2, 2^ge2: gap factors of the scaled sweeps
    int* pst;                // per pair: kPst* state words of the scaled sweeps
    int sel;                 // log-domain affine sweeps: 0 every pair, 1 only flagged pairs
    int* grid;               // grid-scheduling region (null for more than kGridMaxPairs pairs)
    int gap_sum;             // NW / SV reverse sweeps: g0 = opens + extends (linear-gap models)
};

// Claim the next pair for this workgroup (largest first).  Returns -1 when done.
// All threads must call it.  The leading barrier orders the claim after every
// wave's use of the previous one (s_pair, the progress counters).
template <int W>
DEV int claim_pair(const SwArgs& a, int* s_pair, int* prog, int slot) {
    __syncthreads();
    if (tid_x() == 0) *s_pair = atomicAdd(&a.claims[slot], 1);
    for (int q = tid_x(); q < a.max_strips; q += 64 * W) prog[q] = 0;
    __syncthreads();
    // wave-uniform from here on: descriptor loads and buffer resources derived
    // from the pair index stay in scalar registers
    const int pi = __builtin_amdgcn_readfirstlane(*s_pair);
    return pi < a.B ? __builtin_amdgcn_readfirstlane(a.order[pi]) : -1;
}

// ==========================================================================
// Affine gap: forward (value; optionally the DP table for the backward)
// ==========================================================================

// Chunks in which every lane's cell is inside the pair (a full strip, past
// the 63-step ramp and before the right edge) run without per-cell masks.
DEV bool interior_fwd(int nrow, int k, int L2) { return nrow == kStrip && k >= 4 && kChunk * k + kChunk <= L2; }

template <bool STORE>
DEV void affine_fwd_strip(const SwArgs& a, const PairDesc& d, int s, int lane, int* prog,
                          float& zref, float& z, float& zc_comp) {
    const int r = lane;
    const int nrow = min(kStrip, d.L1 - s * kStrip);
    const bool row_ok = r < nrow;
    const int L2 = d.L2, NT = L2 + 63;
    const float* srow = a.scores + d.s_off + (int64_t)(s * kStrip + r) * d.ld;
    // checkpoints: the sweep state before every kCkpt-th step
    constexpr int CP = kWave * kPlanesAffineCkpt;
    const int nblk = (NT + kCkpt - 1) / kCkpt;
    float* ck = a.table + d.t_off + (int64_t)s * nblk * CP + r;  // this lane's column
    const bool has_in = s > 0, has_out = s + 1 < d.ns;
    const int64_t blk = (int64_t)L2 * kBndStride;
    const float* bin = a.bnd + d.b_off + (int64_t)(has_in ? s - 1 : 0) * blk;
    float* bout = a.bnd + d.b_off + (int64_t)s * blk;
    const uint32_t vt = (uint32_t)r * 4;
    const float kap = a.kappa, go2 = a.go2, ge2 = a.ge2;

    float Ml = kNinf, Dl = kNinf;             // M, D of (i, j-1)
    float Mp = kNinf, Ip = kNinf, Ap = kNinf;  // own values of the previous step
    float Adg = kNinf;                         // A of (i-1, j-1)
    float zlim = zref + 64.0f;
    float cur[16], nxt[16];
    load16(cur, srow, -r, L2, row_ok);
    const int nchunk = (NT + kChunk - 1) / kChunk;
    for (int k = 0; k < nchunk; ++k) {
        load16(nxt, srow, kChunk * (k + 1) - r, L2, row_ok);
        // boundary row above: entry u (column 16k + u + 1) in lane u
        float qM = kNinf, qI = kNinf, qA = kNinf;
        if (has_in) {
            const int need = chunk_need(k, L2);
            if (need > 0) wait_ready(&prog[s - 1], need);
            if (lane < kChunk && kChunk * k + lane < L2) {
                const uint32_t off = (uint32_t)(kChunk * k + lane) * (kBndStride * 4);
                qM = gload(bin, off); qI = gload(bin, off + 4); qA = gload(bin, off + 8);
            }
        }
        auto body = [&](auto edge) {
            constexpr bool EDGE = decltype(edge)::value;
            float* ck_k = ck + (int64_t)(kChunk / kCkpt) * k * CP;  // block 2k
            float zc = 0.0f;
#pragma unroll
            for (int u = 0; u < kChunk; ++u) {
                const int t = kChunk * k + u;
                const int j = t - r + 1;
                const bool ok = !EDGE || (row_ok && j >= 1 && j <= L2);
                if (STORE && u % kCkpt == 0 && (!EDGE || t < NT)) {
                    // streaming stores: read back only by the reverse sweep
                    float* q = ck_k + (u / kCkpt) * CP;
                    __builtin_nontemporal_store(Mp, q);
                    __builtin_nontemporal_store(Ip, q + kWave);
                    __builtin_nontemporal_store(Dl, q + 2 * kWave);
                    __builtin_nontemporal_store(Ap, q + 3 * kWave);
                    __builtin_nontemporal_store(Adg, q + 4 * kWave);
                }
                const float Mu = shift_in_prev(Mp, qM);
                const float Iu = shift_in_prev(Ip, qI);
                const float Au = shift_in_prev(Ap, qA);
                float M = cur[u] * kap + softplus2(Adg);
                float I = lse2(Mu + go2, Iu + ge2);
                float D = lse2(Ml + go2, Dl + ge2);
                float A = lse3(M, I, D);
                if (EDGE && !ok) { M = kNinf; I = kNinf; D = kNinf; A = kNinf; }
                // partition: rescale when A outgrows the lane's reference by
                // 2^64 (rare, so the wave branches around it on a scalar mask)
                if (any_greater(A, zlim)) {
                    if (A > zlim && ok) {
                        const float f = ex2(zref - A);
                        z *= f; zc *= f; zc_comp *= f;
                        zref = A;
                        zlim = A + 64.0f;
                    }
                }
                const float e = ex2(A - zref);
                zc += (!EDGE || ok) ? e : 0.0f;
                if (has_out && r == kWave - 1 && ok) {
                    float* e = bout + (int64_t)(t - 63) * kBndStride;  // column j - 1 of lane 63
                    e[0] = M; e[1] = I; e[2] = A;
                }
                Adg = Au; Ml = M; Dl = D; Mp = M; Ip = I; Ap = A;
            }
            kadd(z, zc_comp, zc);
        };
        if (interior_fwd(nrow, k, L2)) body(std::false_type{});
        else body(std::true_type{});
        if (has_out) publish(&prog[s], chunk_done(k, L2), lane);
#pragma unroll
        for (int u = 0; u < 16; ++u) cur[u] = nxt[u];
    }
}

// Combine per-lane partition partials of every wave with the empty alignment
// (value 0) into S in log2 units.
DEV float finish_partition(float zref, float z, float* s_red, int W, int lane, int wv) {
    const float m = wave_max(zref);
    const float zz = wave_sum(z * ex2(zref - m));
    if (lane == 0) { s_red[2 * wv] = m; s_red[2 * wv + 1] = zz; }
    __syncthreads();
    float St = 0.0f;
    if (tid_x() == 0) {
        float M = 0.0f;
        for (int w = 0; w < W; ++w) M = fmaxf(M, s_red[2 * w]);
        float Z = ex2(-M);
        for (int w = 0; w < W; ++w) Z += s_red[2 * w + 1] * ex2(s_red[2 * w] - M);
        St = M + lg2(Z);
    }
    return St;
}

template <bool STORE, int W>
__global__ __launch_bounds__(W * 64) void sw_affine_forward_kernel(SwArgs a) {
    extern __shared__ int s_prog[];
    __shared__ int s_pair;
    __shared__ float s_red[2 * W];
    const int lane = tid_x() & 63, wv = __builtin_amdgcn_readfirstlane(tid_x() >> 6);
    for (;;) {
        if (a.sel && __builtin_amdgcn_readfirstlane(a.claims[kClaimFlagged]) == 0) return;  // uniform
        const int p = claim_pair<W>(a, &s_pair, s_prog, a.sel ? 4 : 0);
        if (p < 0) break;
        if (a.sel && __builtin_amdgcn_readfirstlane(a.pst[4 * p + kPstFlag]) == 0) continue;
        const PairDesc d = a.desc[p];
        float zref = kNinf, z = 0.0f, zcomp = 0.0f;
        for (int s = wv; s < d.ns; s += W) affine_fwd_strip<STORE>(a, d, s, lane, s_prog, zref, z, zcomp);
        const float St = finish_partition(zref, z, s_red, W, lane, wv);
        if (tid_x() == 0) {
            a.stil[p] = St;
            a.value[p] = St / a.kappa;
        }
        __syncthreads();
    }
}

// ==========================================================================
// Affine gap: backward (map and parameter gradients), pull-form adjoint.
// Strips run bottom-up, columns right to left.
// ==========================================================================

// ==========================================================================
// Reverse sweep with recomputation.  The forward sweep kept its state every
// kCkpt = 8 steps; the reverse sweep walks the strip's 8-step blocks from the
// last to the first and for each block
//   1. recomputes the block's M, I, D (registers) from its checkpoint, its
//      scores and the row above (the strip above's forward boundary),
//   2. runs the pull-form adjoint over the block's steps in reverse order.
// The block's scores serve both the recomputation and sum(P * s).  A block's
// inputs are loaded when the block starts; the other waves on the SIMD cover
// the latency (software prefetch one block ahead cost ~55 VGPRs, halving
// occupancy).  Strip hand-off (row below -> this strip's lane 63) works in forward
// 16-step chunks processed in reverse: after chunk k a strip has finished
// every column >= 16k + 1 of its first row.
// ==========================================================================

DEV int rev_need(int k, int L2) {
    // columns read by lane 63 in chunk k: [16k - 62, 16k - 47] within [1, L2]
    const int lo = max(1, kChunk * k - 62), hi = min(L2, kChunk * k - 47);
    return lo > hi ? 0 : L2 - lo + 1;
}
DEV int rev_done(int k, int L2) { return max(0, min(L2, L2 - kChunk * k)); }

// Inputs of one 8-step block.
struct BlockIn {
    float c[kPlanesAffineCkpt];  // checkpoint: M, I, D, A of the previous step, A of the diagonal
    float sc[kCkpt];             // scores of the block's columns (sc[v]: step 8b + v)
    float qM, qI, qA;            // row above, entry v (column 8b + v + 1) in lane v
};

// Map staging: the reverse sweep writes each cell's map value into a per-wave
// LDS ring [32 steps][64 lanes] (slot = step mod 32, the same for every lane,
// so the writes use constant offsets and hit 64 distinct banks) and stores a
// lane's row to global memory in aligned 16-column segments once a segment is
// complete (after the chunk that processed its lowest column).  Direct
// per-lane stores would write 32 misaligned bytes of 64 different rows per
// block, which the memory system turns into repeated partial-line writes.
constexpr int kRing = 32;

// Store lane r's staged segment [cs, cs + 16) of its map row (columns
// outside [0, L2) are skipped).  Column c of lane r was computed at step c + r.
DEV void flush_segment(const float* ring, int r, float* mrow, int cs, int L2) {
    const int t0 = cs + r;
    if (cs >= 0 && cs + 16 <= L2) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            f4u v;
            v.x = ring[((t0 + 4 * q + 0) & (kRing - 1)) * kWave + r];
            v.y = ring[((t0 + 4 * q + 1) & (kRing - 1)) * kWave + r];
            v.z = ring[((t0 + 4 * q + 2) & (kRing - 1)) * kWave + r];
            v.w = ring[((t0 + 4 * q + 3) & (kRing - 1)) * kWave + r];
            *reinterpret_cast<f4u*>(mrow + cs + 4 * q) = v;
            // one group of four in flight at a time (hoisting all sixteen
            // reads with their addresses costs ~30 registers)
            asm volatile("" ::: "memory");
        }
    } else {
        const int lo = max(0, -cs), hi = min(16, L2 - cs);
#pragma unroll 1
        for (int u = lo; u < hi; ++u) mrow[cs + u] = ring[((t0 + u) & (kRing - 1)) * kWave + r];
    }
}

DEV void affine_bwd_strip(const SwArgs& a, const PairDesc& d, int s, int lane, int* prog, float St,
                          float* ring, float (&acc)[3], float (&comp)[3]) {
    const int r = lane;
    const int nrow = min(kStrip, d.L1 - s * kStrip);
    const bool row_ok = r < nrow;
    const int L2 = d.L2, NT = L2 + 63;
    const int nblk = (NT + kCkpt - 1) / kCkpt;
    const int nchunk = (NT + kChunk - 1) / kChunk;
    const float* srow = a.scores + d.s_off + (int64_t)(s * kStrip + r) * d.ld;
    float* mrow = a.map ? a.map + d.s_off + (int64_t)(s * kStrip + r) * d.ld : nullptr;
    const int al = (int)((d.s_off + (int64_t)(s * kStrip + r) * d.ld) & 15);  // row start mod 16 floats
    constexpr int CP = kWave * kPlanesAffineCkpt;
    const float* ck = a.table + d.t_off + (int64_t)s * nblk * CP + r;
    const bool has_up = s > 0, has_in = s + 1 < d.ns, has_out = s > 0;
    const int64_t blk = (int64_t)L2 * kBndStride;
    const float* bup = a.bnd + d.b_off + (int64_t)(has_up ? s - 1 : 0) * blk;  // [0..2]: M, I, A
    const float* bin = a.bnd + d.b_off + (int64_t)(has_in ? s + 1 : s) * blk;  // [4..6]: bI, I, bM
    float* bout = a.bnd + d.b_off + (int64_t)s * blk;
    const float kap = a.kappa, go2 = a.go2, ge2 = a.ge2;

    auto load_block = [&](BlockIn& in, int b) {
        const int t0 = kCkpt * b;
        if (t0 < NT) {
            const float* q = ck + (int64_t)b * CP;
#pragma unroll
            for (int p = 0; p < kPlanesAffineCkpt; ++p) in.c[p] = q[p * kWave];
        } else {
#pragma unroll
            for (int p = 0; p < kPlanesAffineCkpt; ++p) in.c[p] = kNinf;
        }
        load_run<kCkpt>(in.sc, srow, t0 - r, L2, row_ok);
        in.qM = kNinf; in.qI = kNinf; in.qA = kNinf;
        if (has_up && lane < kCkpt && t0 + lane < L2) {
            const float* e = bup + (int64_t)(t0 + lane) * kBndStride;
            in.qM = e[0]; in.qI = e[1]; in.qA = e[2];
        }
    };

    float bDs = 0.0f, Ds = kNinf;              // beta_D, D of (i, j+1)
    float bIp = 0.0f, Ip = kNinf, bMp = 0.0f;  // own values of the previous (later) step
    float bMdg = 0.0f;                         // beta_M of (i+1, j+1)
    for (int k = nchunk - 1; k >= 0; --k) {
        // row below (its lane 0's reverse boundary): entry e (column
        // 16k - 47 - e, processing order) in lane 63 - e
        float qB = 0.0f, qIs = kNinf, qBM = 0.0f;
        if (has_in) {
            const int need = rev_need(k, L2);
            if (need > 0) wait_ready(&prog[s + 1], need);
            const int e = kWave - 1 - lane;
            const int j = kChunk * k - 47 - e;
            if (e < kChunk && j >= 1 && j <= L2) {
                const float* q = bin + (int64_t)(j - 1) * kBndStride;
                qB = q[4]; qIs = q[5]; qBM = q[6];
            }
        }
        float cacc[3] = {0.0f, 0.0f, 0.0f};
        auto body = [&](auto edge) {
            constexpr bool EDGE = decltype(edge)::value;
#pragma unroll
            for (int h = kChunk / kCkpt - 1; h >= 0; --h) {
                BlockIn in;
                const int b = (kChunk / kCkpt) * k + h;
                load_block(in, b);
                // 1. recompute the block (forward order)
                float rm[kCkpt], ri[kCkpt], rd[kCkpt];
                {
                    float Mp = in.c[0], Ip_ = in.c[1], Dl = in.c[2], Ap = in.c[3], Adg = in.c[4];
                    float qM = in.qM, qI = in.qI, qA = in.qA;
#pragma unroll
                    for (int v = 0; v < kCkpt; ++v) {
                        const int t = kCkpt * b + v;
                        const int j = t - r + 1;
                        const bool ok = !EDGE || (row_ok && j >= 1 && j <= L2);
                        const float Mu = shift_in_prev(Mp, qM);
                        const float Iu = shift_in_prev(Ip_, qI);
                        const float Au = shift_in_prev(Ap, qA);
                        float M = in.sc[v] * kap + softplus2(Adg);
                        float I = lse2(Mu + go2, Iu + ge2);
                        float D = lse2(Mp + go2, Dl + ge2);
                        float A = lse3(M, I, D);
                        if (EDGE && !ok) { M = kNinf; I = kNinf; D = kNinf; A = kNinf; }
                        rm[v] = M; ri[v] = I; rd[v] = D;
                        Adg = Au; Mp = M; Ip_ = I; Dl = D; Ap = A;
                    }
                }
                // 2. adjoint (reverse order)
#pragma unroll
                for (int v = kCkpt - 1; v >= 0; --v) {
                    const int t = kCkpt * b + v;
                    const int j = t - r + 1;
                    const bool ok = !EDGE || (row_ok && j >= 1 && j <= L2);
                    const float M = rm[v], I = ri[v], D = rd[v];
                    const float bIs = shift_in_next(bIp, qB);
                    const float Is = shift_in_next(Ip, qIs);
                    const float bMs = shift_in_next(bMp, qBM);
                    float A, wM, wI, wD;
                    lse3_norm(M, I, D, A, wM, wI, wD);
                    const float bA = ex2(fminf(A - St, 0.0f)) + bMdg * sigmoid2(A);
                    const float mg = M + go2;
                    float wIM, wII, wDM, wDD;
                    pair_weights(mg - Is, I + ge2 - Is, wIM, wII);
                    pair_weights(mg - Ds, D + ge2 - Ds, wDM, wDD);
                    float tIM = bIs * wIM, tII = bIs * wII, tDM = bDs * wDM, tDD = bDs * wDD;
                    float tMo = tIM + tDM, tEx = tII + tDD;
                    float bM = bA * wM + tMo;
                    float bI = bA * wI + tII;
                    float bD = bA * wD + tDD;
                    if (EDGE && !ok) { bM = 0.0f; bI = 0.0f; bD = 0.0f; tMo = 0.0f; tEx = 0.0f; }
                    if (!EDGE || ok) ring[(t & (kRing - 1)) * kWave + r] = bM;
                    cacc[0] += tMo;
                    cacc[1] += tEx;
                    cacc[2] += bM * in.sc[v];
                    // keep the running sums in step order (the scheduler would
                    // otherwise defer the adds and hold every summand live)
                    asm volatile("" : "+v"(cacc[0]), "+v"(cacc[1]), "+v"(cacc[2]));
                    if (has_out && r == 0 && ok) {
                        float* e = bout + (int64_t)t * kBndStride;  // column j - 1 of lane 0
                        e[4] = bI; e[5] = I; e[6] = bM;
                    }
                    bMdg = bMs; bDs = bD; Ds = D; bIp = bI; Ip = I; bMp = bM;
                }
            }
        };
        if (interior_fwd(nrow, k, L2)) body(std::false_type{});
        else body(std::true_type{});
        // Segments are 16 floats aligned in memory (a segment starts at column
        // cs with (cs + al) % 16 == 0); the row's first and last may be
        // partial.  Each chunk covers 16 columns of the lane's row, so exactly
        // one segment start falls in it; that segment is complete once the
        // chunk is done (higher columns were processed by earlier chunks).
        // The row's first segment, when it starts below column 0, is stored
        // after the strip.  Unstored columns span fewer than 32 steps, so the
        // 32-slot ring is not overwritten before its columns are stored.
        if (mrow && row_ok) {
            const int cl = kChunk * k - r, ch = cl + kChunk - 1;
            const int cb = ((max(cl, 0) + al + 15) & ~15) - al;  // first start >= max(cl, 0)
            if (cb <= ch && cb < L2) flush_segment(ring, r, mrow, cb, L2);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) kadd(acc[c], comp[c], cacc[c]);
        if (has_out) publish(&prog[s], rev_done(k, L2), lane);
    }
    if (mrow && row_ok && al != 0 && L2 > 0) flush_segment(ring, r, mrow, -al, L2);
}

// Fixed-order sum of per-lane accumulators over the workgroup.
template <int N>
DEV void block_sums(float (&acc)[N], float* s_red, int W, int lane, int wv, float (&out)[N]) {
#pragma unroll
    for (int c = 0; c < N; ++c) {
        float v = wave_sum(acc[c]);
        if (lane == 0) s_red[N * wv + c] = v;
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < N; ++c) {
        float t = 0.0f;
        for (int w = 0; w < W; ++w) t += s_red[N * w + c];
        out[c] = t;
    }
}

template <int W>
__global__ __launch_bounds__(W * 64) void sw_affine_backward_kernel(SwArgs a) {
    extern __shared__ int s_prog[];
    __shared__ int s_pair;
    __shared__ float s_red[3 * W];
    __shared__ float s_ring[W * kRing * kWave];
    const int lane = tid_x() & 63, wv = __builtin_amdgcn_readfirstlane(tid_x() >> 6);
    for (;;) {
        if (a.sel && __builtin_amdgcn_readfirstlane(a.claims[kClaimFlagged]) == 0) return;  // uniform
        const int p = claim_pair<W>(a, &s_pair, s_prog, a.sel ? 5 : 1);
        if (p < 0) break;
        if (a.sel && __builtin_amdgcn_readfirstlane(a.pst[4 * p + kPstFlag]) == 0) continue;
        const PairDesc d = a.desc[p];
        const float St = a.stil[p];
        float acc[3] = {0.0f, 0.0f, 0.0f}, comp[3] = {0.0f, 0.0f, 0.0f};
        // bottom-up: the q-th strip of the pair in processing order is ns-1-q
        for (int q = wv; q < d.ns; q += W)
            affine_bwd_strip(a, d, d.ns - 1 - q, lane, s_prog, St, s_ring + wv * (kRing * kWave), acc, comp);
        float tot[3];
        block_sums<3>(acc, s_red, W, lane, wv, tot);
        if (tid_x() == 0) {
            const float S = a.value[p];
            if (a.g0) a.g0[p] = tot[0];
            if (a.g1) a.g1[p] = tot[1];
            if (a.g2) a.g2[p] = (S - (tot[2] + a.go * tot[0] + a.ge * tot[1])) / a.temp;
        }
        if (a.padded) zero_padding(a.map, d, a.L1max, a.L2max, tid_x(), 64 * W);
        __syncthreads();
    }
}

#include "nw_affine.hiph"
#include "nw_tangent.hiph"
#include "sv_affine.hiph"
#include "sv_tangent.hiph"
#include "sw_teams.hiph"
#include "sw_grid.hiph"
#include "sw_affine_scaled.hiph"
#include "sw_linear_scaled.hiph"
#include "embed_fused.hiph"

// ==========================================================================
// Linear gap: forward.  Table planes: A (cell value), X = s + LSE(A', 0).
// ==========================================================================

template <bool STORE>
DEV void linear_fwd_strip(const SwArgs& a, const PairDesc& d, int s, int lane, int* prog,
                          float& zref, float& z, float& zc_comp) {
    const int r = lane;
    const int i = s * kStrip + r + 1;
    const bool row_ok = i <= d.L1;
    const int L2 = d.L2, NT = L2 + 63;
    const float* srow = a.scores + d.s_off + (int64_t)(i - 1) * d.ld;
    float* tab = a.table + d.t_off + (int64_t)s * NT * (kWave * kPlanesLinearFwd);
    const float* bin = s > 0 ? a.bnd + d.b_off + (int64_t)(s - 1) * L2 * kBndStride : nullptr;
    float* bout = s + 1 < d.ns ? a.bnd + d.b_off + (int64_t)s * L2 * kBndStride : nullptr;
    const float kap = a.kappa, g2 = a.go2;

    float Al = kNinf, Ap = kNinf, Adg = kNinf;
    float cur[16], nxt[16];
    load16(cur, srow, -r, L2, row_ok);
    const int nchunk = (NT + kChunk - 1) / kChunk;
    for (int k = 0; k < nchunk; ++k) {
        load16(nxt, srow, kChunk * (k + 1) - r, L2, row_ok);
        float eA = kNinf;
        if (bin) {
            const int need = chunk_need(k, L2);
            if (need > 0) wait_ready(&prog[s - 1], need);
            const int j = kChunk * k + lane + 1;
            if (lane < kChunk && j <= L2) eA = bin[(int64_t)(j - 1) * kBndStride];
        }
        float zc = 0.0f;
#pragma unroll
        for (int u = 0; u < kChunk; ++u) {
            const int t = kChunk * k + u;
            const int j = t - r + 1;
            const bool ok = row_ok && j >= 1 && j <= L2;
            const float Au = from_prev_lane(Ap, lane_value(eA, u));
            float X = cur[u] * kap + softplus2(Adg);
            float A = lse3(X, Au + g2, Al + g2);
            if (!ok) { X = kNinf; A = kNinf; }
            if (ok && A > zref + 64.0f) {
                const float f = ex2(zref - A);
                z *= f; zc *= f; zc_comp *= f;
                zref = A;
            }
            zc += ok ? ex2(A - zref) : 0.0f;
            if (STORE && t < NT) {
                float* q = tab + (int64_t)t * (kWave * kPlanesLinearFwd) + r;
                q[0] = A; q[kWave] = X;
            }
            if (bout && r == kWave - 1 && ok) bout[(int64_t)(j - 1) * kBndStride] = A;
            Adg = Au; Al = A; Ap = A;
        }
        kadd(z, zc_comp, zc);
        if (bout) publish(&prog[s], chunk_done(k, L2), lane);
#pragma unroll
        for (int u = 0; u < 16; ++u) cur[u] = nxt[u];
    }
}

template <bool STORE, int W>
__global__ __launch_bounds__(W * 64) void sw_forward_kernel(SwArgs a) {
    extern __shared__ int s_prog[];
    __shared__ int s_pair;
    __shared__ float s_red[2 * W];
    const int lane = tid_x() & 63, wv = __builtin_amdgcn_readfirstlane(tid_x() >> 6);
    for (;;) {
        if (a.sel && __builtin_amdgcn_readfirstlane(a.claims[kClaimFlagged]) == 0) return;  // uniform
        const int p = claim_pair<W>(a, &s_pair, s_prog, a.sel ? 4 : 0);
        if (p < 0) break;
        if (a.sel && __builtin_amdgcn_readfirstlane(a.pst[4 * p + kPstFlag]) == 0) continue;
        const PairDesc d = a.desc[p];
        float zref = kNinf, z = 0.0f, zcomp = 0.0f;
        for (int s = wv; s < d.ns; s += W) linear_fwd_strip<STORE>(a, d, s, lane, s_prog, zref, z, zcomp);
        const float St = finish_partition(zref, z, s_red, W, lane, wv);
        if (tid_x() == 0) {
            a.stil[p] = St;
            a.value[p] = St / a.kappa;
        }
        __syncthreads();
    }
}

// ==========================================================================
// Linear gap: backward, edge-flow form.  Node c = (i, j) splits its adjoint
// beta(c) = 2^(A-S) + gamma(i+1,j+1) sig(A) + fu(i+1,j) + fl(i,j+1)
// over its in-edges with weights normalised at c:
//   gamma(c) = beta wx (to X, = map), fu(c) = beta wu (from (i-1, j)),
//   fl(c) = beta wl (from (i, j-1)).
// A(i, j-1) is the table entry one step earlier in the same lane; A(i-1, j)
// is that entry of lane r-1 (DPP), or the strip above's forward boundary.
// Boundary entries: forward A at [0]; backward fu, gamma at [4], [5].
// ==========================================================================

DEV void linear_bwd_strip(const SwArgs& a, const PairDesc& d, int s, int lane, int* prog, float St,
                          float (&acc)[2], float (&comp)[2]) {
    const int r = lane;
    const int i = s * kStrip + r + 1;
    const bool row_ok = i <= d.L1;
    const int L2 = d.L2, NT = L2 + 63;
    const float* srow = a.scores + d.s_off + (int64_t)(i - 1) * d.ld;
    float* mrow = a.map ? a.map + d.s_off + (int64_t)(i - 1) * d.ld : nullptr;
    const float* tab = a.table + d.t_off + (int64_t)s * NT * (kWave * kPlanesLinearFwd);
    const float* bup = s > 0 ? a.bnd + d.b_off + (int64_t)(s - 1) * L2 * kBndStride : nullptr;
    const float* bin = s + 1 < d.ns ? a.bnd + d.b_off + (int64_t)(s + 1) * L2 * kBndStride : nullptr;
    float* bout = s > 0 ? a.bnd + d.b_off + (int64_t)s * L2 * kBndStride : nullptr;
    const float g2 = a.go2;
    constexpr int P = kWave * kPlanesLinearFwd;

    float flP = 0.0f;              // fl of (i, j+1)
    float fuP = 0.0f, gP = 0.0f;   // own fu, gamma of the previous step
    float gDg = 0.0f;              // gamma of (i+1, j+1)
    float Acur = tab[(int64_t)(NT - 1) * P + r];
    float cur[16], nxt[16];
    load16(cur, srow, NT - kChunk - r, L2, row_ok);
    const int nchunk = (NT + kChunk - 1) / kChunk;
    for (int k = 0; k < nchunk; ++k) {
        load16(nxt, srow, NT - kChunk * (k + 2) - r, L2, row_ok);
        float eF = 0.0f, eG = 0.0f, eA = kNinf;
        if (bup) {
            const int idx = L2 + 62 - (kChunk * k + lane);  // lane 0's column - 1 at step u = lane
            if (lane < kChunk && idx >= 0 && idx < L2) eA = bup[(int64_t)idx * kBndStride];
        }
        if (bin) {
            const int need = chunk_need(k, L2);
            if (need > 0) wait_ready(&prog[s + 1], need);
            const int j = L2 - (kChunk * k + lane);
            if (lane < kChunk && j >= 1) {
                const float* e = bin + (int64_t)(j - 1) * kBndStride;
                eF = e[4]; eG = e[5];
            }
        }
        float pc[16];
        float cacc[2] = {0.0f, 0.0f};
#pragma unroll
        for (int u = 0; u < kChunk; ++u) {
            const int t = NT - 1 - (kChunk * k + u);
            const int j = t - r + 1;
            const bool ok = row_ok && j >= 1 && j <= L2;
            const float Aleft = t >= 1 ? tab[(int64_t)(t - 1) * P + r] : kNinf;
            const float X = tab[(int64_t)max(t, 0) * P + kWave + r];
            const float Aup = from_prev_lane(Aleft, lane_value(eA, u));
            const float inU = from_next_lane(fuP, lane_value(eF, u));
            const float gD = from_next_lane(gP, lane_value(eG, u));
            float y, wx, wu, wl;
            lse3_norm(X, Aup + g2, Aleft + g2, y, wx, wu, wl);
            const float bt = ex2(fminf(Acur - St, 0.0f)) + gDg * sigmoid2(Acur) + inU + flP;
            float gm = bt * wx, fu = bt * wu, fl = bt * wl;
            if (!ok) { gm = 0.0f; fu = 0.0f; fl = 0.0f; }
            pc[u] = gm;
            cacc[0] += fu + fl;
            cacc[1] += gm * cur[kChunk - 1 - u];
            if (bout && r == 0 && ok) {
                float* e = bout + (int64_t)(j - 1) * kBndStride;
                e[4] = fu; e[5] = gm;
            }
            gDg = gD; flP = fl; fuP = fu; gP = gm; Acur = Aleft;
        }
#pragma unroll
        for (int c = 0; c < 2; ++c) kadd(acc[c], comp[c], cacc[c]);
        if (mrow) {
            float outv[16];
#pragma unroll
            for (int v = 0; v < 16; ++v) outv[v] = pc[kChunk - 1 - v];
            store16(mrow, NT - kChunk * (k + 1) - r, L2, row_ok, outv);
        }
        if (bout) publish(&prog[s], chunk_done(k, L2), lane);
#pragma unroll
        for (int u = 0; u < 16; ++u) cur[u] = nxt[u];
    }
}

template <int W>
__global__ __launch_bounds__(W * 64) void sw_backward_kernel(SwArgs a) {
    extern __shared__ int s_prog[];
    __shared__ int s_pair;
    __shared__ float s_red[2 * W];
    const int lane = tid_x() & 63, wv = __builtin_amdgcn_readfirstlane(tid_x() >> 6);
    for (;;) {
        if (a.sel && __builtin_amdgcn_readfirstlane(a.claims[kClaimFlagged]) == 0) return;  // uniform
        const int p = claim_pair<W>(a, &s_pair, s_prog, a.sel ? 5 : 1);
        if (p < 0) break;
        if (a.sel && __builtin_amdgcn_readfirstlane(a.pst[4 * p + kPstFlag]) == 0) continue;
        const PairDesc d = a.desc[p];
        const float St = a.stil[p];
        float acc[2] = {0.0f, 0.0f}, comp[2] = {0.0f, 0.0f};
        for (int q = wv; q < d.ns; q += W)
            linear_bwd_strip(a, d, d.ns - 1 - q, lane, s_prog, St, acc, comp);
        float tot[2];
        block_sums<2>(acc, s_red, W, lane, wv, tot);
        if (tid_x() == 0) {
            const float S = a.value[p];
            if (a.g0) a.g0[p] = tot[0];
            if (a.g1) a.g1[p] = (S - (tot[1] + a.go * tot[0])) / a.temp;
        }
        if (a.padded) zero_padding(a.map, d, a.L1max, a.L2max, tid_x(), 64 * W);
        __syncthreads();
    }
}

#include "sw_tangent.hiph"
#include "sw_tangent_scaled.hiph"


// ==========================================================================
// Host side: argument checks, launch geometry, C ABI.
// ==========================================================================

namespace {

ohm_status fail(ohm_status st, const std::string& msg) {
    set_error(msg);
    return st;
}

ohm_status check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(OHM_ELAUNCH, std::string(what) + ": " + hipGetErrorString(e));
    return OHM_OK;
}

// Tail of every packed-layout GPU call: lost_pairs_kernel on the claims and
// descriptors that setup() laid out at the front of the workspace.  (Padded
// calls need none: setup() requires the padded bound, so every pair fits.)
ohm_status finish_lost(const int32_t* lengths, const int64_t* offsets, int B, int L1max,
                       int L2max, void* ws, const CallOut& o, void* stream) {
    if (!offsets || B == 0) return OHM_OK;
    size_t claims, desc, order, stil, tgt, pst, grid;
    ws_fixed(B, L1max > 0 ? (L1max + kStrip - 1) / kStrip : 1, &claims, &desc, &order, &stil, &tgt,
             &pst, &grid);
    char* base = static_cast<char*>(ws);
    hipLaunchKernelGGL(lost_pairs_kernel, dim3(1), dim3(256), 0, static_cast<hipStream_t>(stream),
                       reinterpret_cast<const int*>(base + claims),
                       reinterpret_cast<const PairDesc*>(base + desc), lengths, offsets, B, L1max,
                       L2max, o);
    return check_launch("lost-pairs kernel");
}

int device_cus() {
    static std::mutex mu;
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    std::lock_guard<std::mutex> lock(mu);
    if (!cached[dev]) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            n <= 0)
            n = 256;
        cached[dev] = n;
    }
    return cached[dev];
}

int env_int(const char* name, int fallback) {
    const char* e = getenv(name);
    return e ? atoi(e) : fallback;
}

// Launch geometry.  Waves per workgroup W: more waves shorten one pair's
// critical path (its strips run concurrently), fewer waste less on short
// pairs.  Workgroups: enough persistent groups for 16 waves per CU, never more
// than there are pairs.  OHM_WAVES / OHM_WG_PER_CU override for tuning.
// Optional per-launch timing events (ohm_set_stage_events).
constexpr int kStageMarks = 5;
void* g_stage_events[kStageMarks] = {};

void stage_mark(int k, hipStream_t stream) {
    if (k < 0 || k >= kStageMarks) return;
    if (g_stage_events[k]) (void)hipEventRecord(static_cast<hipEvent_t>(g_stage_events[k]), stream);
}

struct Geometry {
    int waves, groups;
};

// Launch configuration set through ohm_set_launch_config (0 = automatic);
// it takes precedence over the OHM_* environment overrides.
std::atomic<int> g_cfg_waves{0}, g_cfg_per_cu{0}, g_cfg_teams{0};
// Checkpoint stride of the scaled affine sweeps (ohm_set_checkpoint_stride;
// 0 = automatic: kCkpt).
std::atomic<int> g_cfg_ckpt{0};

int ckpt_stride() {
    const int v = g_cfg_ckpt.load();
    return v ? v : kCkpt;
}
// Table stride the plan kernel lays out for the affine forward + reverse
// sweeps: the scaled sweeps' stride, or the log-domain sweeps' kCkpt (which
// redo flagged pairs in the same table) when that is finer.
int affine_table_stride(int ck) { return ck < kCkpt ? ck : kCkpt; }

Geometry choose_geometry(int B, int L1max) {
    const int cus = device_cus();
    const int strips = (L1max + kStrip - 1) / kStrip;
    int waves = (B >= 2 * cus) ? 4 : 8;
    while (waves > 1 && waves > strips) waves >>= 1;
    int v = g_cfg_waves.load() ? g_cfg_waves.load() : env_int("OHM_WAVES", 0);
    if (v == 1 || v == 2 || v == 4 || v == 8) waves = v;
    int per_cu = 16 / waves;
    v = g_cfg_per_cu.load() ? g_cfg_per_cu.load() : env_int("OHM_WG_PER_CU", 0);
    if (v >= 1 && v <= 64) per_cu = v;
    long groups = (long)cus * per_cu;
    if (groups > B) groups = B;
    if (groups < 1) groups = 1;
    return {waves, (int)groups};
}

struct Launch {
    SwArgs a;
    Geometry geo;
    size_t lds;
    hipStream_t stream;
    int ck;  // checkpoint stride of the scaled affine sweeps, read once per call
};

// Operations for setup(): which table the plan kernel lays out.
constexpr int kOpValue = 0;      // value only (no table)
constexpr int kOpForward = 1;    // Smith-Waterman forward + reverse
constexpr int kOpTangent = 2;    // Smith-Waterman tangent passes
constexpr int kOpNw = 3;         // Needleman-Wunsch forward + reverse
constexpr int kOpSv = 4;         // Saigo-Vert forward + reverse
constexpr int kOpNwTangent = 5;  // Needleman-Wunsch tangent passes
constexpr int kOpSvTangent = 6;  // Saigo-Vert tangent passes

ohm_status setup(int op, bool affine, const float* scores, const int32_t* lengths,
                 const int64_t* offsets, int B, int L1max, int L2max, float go, float ge,
                 float temp, void* ws, size_t ws_bytes, void* stream, Launch& L) {
    if (B < 0 || L1max < 0 || L2max < 0) return fail(OHM_EINVAL, "negative batch size or length");
    if (!(temp > 0.0f) || !std::isfinite(temp))
        return fail(OHM_EINVAL, "temperature must be finite and strictly positive");
    if (!std::isfinite(go) || !std::isfinite(ge)) return fail(OHM_EINVAL, "gap scores must be finite");
    // a batch with an empty dimension has no cells: its scores may be null
    if (B > 0 && L1max > 0 && L2max > 0 && !scores) return fail(OHM_EINVAL, "scores is null");
    if (offsets && !lengths) return fail(OHM_EINVAL, "the packed layout needs lengths");
    if ((int64_t)((L1max + kStrip - 1) / kStrip) * 4 > 60 * 1024)
        return fail(OHM_EINVAL, "L1 too long for one launch (max 983040 rows)");
    L.ck = ckpt_stride();
    // the Needleman-Wunsch and Saigo-Vert passes keep full tables
    const TableSpec tf = op == kOpNw ? TableSpec{kPlanesNw, 1}
                         : op == kOpSv ? TableSpec{kPlanesSv, 1}
                         : op == kOpNwTangent ? TableSpec{kPlanesNwTan, 1}
                         : op == kOpSvTangent ? TableSpec{kPlanesSvTan, 1}
                         : affine ? TableSpec{kPlanesAffineCkpt, affine_table_stride(L.ck)}
                                  : TableSpec{kPlanesLinearFwd, 1};
    const TableSpec tt = {affine ? kPlanesAffineTan : kPlanesLinearTan, 1};
    WsLayout w = ws_layout(op >= kOpNw ? kOpForward : op, B, L1max, L2max, -1, tf, tt);
    if (!ws) return fail(OHM_EWORKSPACE, "workspace is null");
    if (ws_bytes < w.data + 256) return fail(OHM_EWORKSPACE, "workspace too small");
    if (!offsets && ws_bytes < w.total)
        return fail(OHM_EWORKSPACE, "workspace smaller than the size query returned");
    char* base = static_cast<char*>(ws);
    SwArgs& a = L.a;
    std::memset(&a, 0, sizeof(a));
    a.scores = scores;
    a.desc = reinterpret_cast<const PairDesc*>(base + w.desc);
    a.order = reinterpret_cast<const int32_t*>(base + w.order);
    a.claims = reinterpret_cast<int*>(base + w.claims);
    a.stil = reinterpret_cast<float*>(base + w.stil);
    a.pst = reinterpret_cast<int*>(base + w.pst);
    a.tgt = reinterpret_cast<float*>(base + w.tgt);
    a.table = reinterpret_cast<float*>(base + w.data);
    a.bnd = a.table;  // per-pair offsets (t_off, b_off) index the same region
    a.B = B;
    a.L1max = L1max;
    a.L2max = L2max;
    a.padded = offsets ? 0 : 1;
    a.max_strips = (L1max + kStrip - 1) / kStrip;
    if (a.max_strips < 1) a.max_strips = 1;
    a.kappa = kLog2e / temp;
    a.go2 = go * a.kappa;
    a.ge2 = ge * a.kappa;
    a.go = go;
    a.ge = ge;
    a.temp = temp;
    a.lgo = std::exp2(a.go2);
    a.lge = std::exp2(a.ge2);
    L.geo = choose_geometry(B, L1max);
    L.lds = (size_t)a.max_strips * sizeof(int);
    L.stream = static_cast<hipStream_t>(stream);
    const TableSpec ts = op == kOpValue ? TableSpec{0, 1} : (op == kOpTangent ? tt : tf);
    const int64_t cap = (int64_t)((ws_bytes - w.data) / sizeof(float));
    if (B <= kGridMaxPairs) a.grid = reinterpret_cast<int*>(base + w.grid);
    hipLaunchKernelGGL(sw_plan_kernel, dim3(1), dim3(kPlanThreads), 0, L.stream, lengths, offsets,
                       B, L1max, L2max, ts.planes, ts.stride, cap, const_cast<PairDesc*>(a.desc),
                       const_cast<int32_t*>(a.order), a.claims, a.grid,
                       a.grid ? (int)grid_zero_ints(B, a.max_strips) : 0);
    stage_mark(0, L.stream);
    return check_launch("plan kernel");
}

// Runs BODY with `constexpr int CK` set to the checkpoint stride CKV (4, 8
// or 16).
#define OHM_WITH_CK(CKV, BODY)                                                            \
    do {                                                                                  \
        switch (CKV) {                                                                    \
            case 4: { constexpr int CK = 4; BODY; break; }                                \
            case 16: { constexpr int CK = 16; BODY; break; }                              \
            default: { constexpr int CK = 8; BODY; break; }                               \
        }                                                                                 \
    } while (0)

#define OHM_DISPATCH_W(L, KERNEL_EXPR)                                                    \
    do {                                                                                  \
        const dim3 grid_((L).geo.groups);                                                 \
        switch ((L).geo.waves) {                                                          \
            case 1: { constexpr int W = 1; hipLaunchKernelGGL(KERNEL_EXPR, grid_, dim3(64 * W), (L).lds, (L).stream, (L).a); break; } \
            case 2: { constexpr int W = 2; hipLaunchKernelGGL(KERNEL_EXPR, grid_, dim3(64 * W), (L).lds, (L).stream, (L).a); break; } \
            case 4: { constexpr int W = 4; hipLaunchKernelGGL(KERNEL_EXPR, grid_, dim3(64 * W), (L).lds, (L).stream, (L).a); break; } \
            default: { constexpr int W = 8; hipLaunchKernelGGL(KERNEL_EXPR, grid_, dim3(64 * W), (L).lds, (L).stream, (L).a); break; } \
        }                                                                                 \
    } while (0)

// Team kernels (wave-level strip scheduling, one workgroup of kTeamWaves
// waves per CU) pay off for batches of few large pairs (measured: 4096-long
// pairs, 16 per batch, 10% faster); for batches with at least a quarter as
// many pairs as CUs the per-pair workgroup kernels are faster (256 x 512^2
// and ragged batches: 15-20%).  They hold per-slot progress counters and partials in
// LDS, and the reverse sweep its waves' map rings; past that size (very tall
// pairs) the workgroup kernels run.  OHM_TEAMS=0/1 forces the choice.
constexpr size_t kLdsLimit = 160 * 1024;
size_t team_bwd_lds_bytes(int max_strips) {
    return team_lds_bytes(max_strips) + (size_t)kTeamWaves * kRing * kWave * sizeof(float);
}
bool use_teams(int B, int max_strips) {
    const int cfg = g_cfg_teams.load();
    const int forced = cfg == 1 ? 1 : cfg == 2 ? 0 : env_int("OHM_TEAMS", -1);
    if (forced == 0) return false;
    if (team_bwd_lds_bytes(max_strips) + sizeof(TeamLds) + 1024 > kLdsLimit) return false;
    return forced == 1 || 4 * B < device_cus();
}
template <typename K>
ohm_status launch_team(K kernel, size_t lds, const Launch& L) {
    if (lds > 64 * 1024 &&
        hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        return fail(OHM_ELAUNCH, "cannot reserve LDS for the team kernel");
    hipLaunchKernelGGL(kernel, dim3(device_cus()), dim3(64 * kTeamWaves), lds, L.stream, L.a);
    return check_launch("team kernel");
}

// Grid kernels (strips of every pair on the waves of all workgroups,
// hand-off through global memory) for batches of few long pairs, which the
// workgroup and team kernels would leave on a few CUs.  teams = 3 in
// ohm_set_launch_config or OHM_GRID=0/1 force the choice.
size_t grid_bwd_lds_bytes(int max_strips) {
    return grid_cum_bytes(max_strips) + (size_t)kGridWaves * kRing * kWave * sizeof(float);
}
// Measured on 16 pairs of 4096 x 4096: grid 6.7 ms, team kernels 21 ms.
bool use_grid(const SwArgs& a) {
    if (!a.grid) return false;
    if (grid_bwd_lds_bytes(a.max_strips) + sizeof(GridLds) + 1024 > kLdsLimit) return false;
    const int cfg = g_cfg_teams.load();
    if (cfg == 1 || cfg == 2 || cfg == 4) return false;  // an explicit other shape
    const int forced = cfg == 3 ? 1 : env_int("OHM_GRID", -1);
    if (forced >= 0) return forced == 1;
    if (env_int("OHM_TEAMS", -1) >= 0) return false;  // an explicit team / workgroup choice
    return 4 * a.B < device_cus() && a.max_strips >= 4;
}
ohm_status launch_grid_fill(const Launch& L) {
    hipLaunchKernelGGL(grid_fill_kernel, dim3(L.a.B, 64), dim3(256), 0, L.stream, L.a);
    return check_launch("grid fill kernel");
}
template <typename K>
ohm_status launch_grid(K kernel, size_t lds, const Launch& L) {
    if (lds > 64 * 1024 &&
        hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        return fail(OHM_ELAUNCH, "cannot reserve LDS for the grid kernel");
    // one wave per strip of the batch, at most 16 waves per CU
    const long want = ((long)L.a.B * L.a.max_strips + kGridWaves - 1) / kGridWaves;
    const long groups = std::max(1L, std::min(want, (long)device_cus() * (16 / kGridWaves)));
    hipLaunchKernelGGL(kernel, dim3((unsigned)groups), dim3(64 * kGridWaves), lds, L.stream, L.a);
    return check_launch("grid kernel");
}

// Fused forward + reverse per pair (sw_affine_scaled_fused_kernel): the
// automatic choice wherever the per-pair workgroup kernels run (measured
// faster than separate launches on 2048 ragged pairs up to 1024^2, 256
// pairs of 512^2 and 512 pairs up to 768^2); teams = 4 / 2 in
// ohm_set_launch_config or OHM_FUSED=1 / 0 force it on / off.
bool use_fused(int B, int max_strips) {
    // its progress words (forward and reverse, 2 x max_strips ints) stay
    // within the default dynamic LDS of a launch; taller batches run
    // separate launches (one set of words each)
    if (2 * (size_t)max_strips * sizeof(int) > 48 * 1024) return false;
    const int cfg = g_cfg_teams.load();
    if (cfg == 4) return true;
    if (cfg != 0) return false;
    const int env = env_int("OHM_FUSED", -1);
    if (env >= 0) return env == 1;
    return !use_teams(B, max_strips);
}

// The scaled affine sweeps need representable gap factors 2^(g log2(e) / T)
// with room for products of two of them and a neighbour ratio; outside that
// (or with OHM_SCALED=0) the log-domain sweeps run alone.
bool use_scaled(const SwArgs& a) {
    if (env_int("OHM_SCALED", 1) == 0) return false;
    return std::fabs(a.go2) <= 120.0f && std::fabs(a.ge2) <= 120.0f;
}

ohm_status forward_impl(bool affine, const float* scores, const int32_t* lengths,
                        const int64_t* offsets, int B, int L1max, int L2max, float go, float ge,
                        float temp, float* value, float* map, float* g0, float* g1, float* g2,
                        void* ws, size_t ws_bytes, void* stream) {
    if (B > 0 && !value) return fail(OHM_EINVAL, "value is null");
    const bool need_bwd = map || g0 || g1 || g2;
    Launch L;
    ohm_status st = setup(need_bwd ? kOpForward : kOpValue, affine, scores, lengths, offsets, B, L1max, L2max, go,
                          ge, temp, ws, ws_bytes, stream, L);
    if (st != OHM_OK || B == 0) return st;
    L.a.value = value;
    L.a.map = map;
    L.a.g0 = g0;
    L.a.g1 = g1;
    L.a.g2 = g2;
    const bool scaled = affine && use_scaled(L.a);
    // linear gap: exponent-tracked sweeps (per-pair workgroups), the
    // log-domain sweeps for the pairs they flag; OHM_LINEAR_SCALED=0 keeps the
    // log-domain sweeps alone
    const bool lin_scaled = !affine && use_scaled(L.a) && env_int("OHM_LINEAR_SCALED", 1) != 0;
    const bool lin_grid = lin_scaled && use_grid(L.a);
    const bool grid = scaled && use_grid(L.a);
    const bool fused = scaled && !grid && need_bwd && use_fused(B, L.a.max_strips);
    const bool team = scaled && !grid && !fused && use_teams(B, L.a.max_strips);
    if (scaled) {
        // scaled sweep for every pair, then the log-domain sweep for the pairs
        // it flagged (a launch whose groups only walk the claim list when
        // none are)
        if (grid) {
            // one fill serves both sweeps (forward entries in words 0..2,
            // reverse entries in words 4..6 of each block)
            if ((st = launch_grid_fill(L)) != OHM_OK) return st;
            const size_t lds = grid_cum_bytes(L.a.max_strips);
            if (need_bwd) OHM_WITH_CK(L.ck, st = launch_grid(sw_affine_scaled_forward_grid<true, CK>, lds, L));
            else st = launch_grid(sw_affine_scaled_forward_grid<false, kCkpt>, lds, L);
            if (st != OHM_OK) return st;
        } else if (fused) {
            Launch F = L;
            F.lds = 2 * (size_t)L.a.max_strips * sizeof(int);
            OHM_WITH_CK(L.ck, OHM_DISPATCH_W(F, (sw_affine_scaled_fused_kernel<W, CK>)));
        } else if (team) {
            const size_t lds = team_lds_bytes(L.a.max_strips);
            if (need_bwd) OHM_WITH_CK(L.ck, st = launch_team(sw_affine_scaled_forward_team<true, CK>, lds, L));
            else st = launch_team(sw_affine_scaled_forward_team<false, kCkpt>, lds, L);
            if (st != OHM_OK) return st;
        } else {
            if (need_bwd) OHM_WITH_CK(L.ck, OHM_DISPATCH_W(L, (sw_affine_scaled_forward_kernel<true, W, CK>)));
            else OHM_DISPATCH_W(L, (sw_affine_scaled_forward_kernel<false, W, kCkpt>));
        }
        stage_mark(1, L.stream);
        L.a.sel = 1;
        if (need_bwd) OHM_DISPATCH_W(L, (sw_affine_forward_kernel<true, W>));
        else OHM_DISPATCH_W(L, (sw_affine_forward_kernel<false, W>));
    } else if (affine) {
        if (need_bwd) OHM_DISPATCH_W(L, (sw_affine_forward_kernel<true, W>));
        else OHM_DISPATCH_W(L, (sw_affine_forward_kernel<false, W>));
    } else if (lin_scaled) {
        if (lin_grid) {
            // one fill serves both sweeps (forward entries in words 0..1,
            // reverse entries in words 4..6 of each block)
            if ((st = launch_grid_fill(L)) != OHM_OK) return st;
            const size_t lds = grid_cum_bytes(L.a.max_strips);
            if (need_bwd) st = launch_grid(sw_linear_scaled_forward_grid<true>, lds, L);
            else st = launch_grid(sw_linear_scaled_forward_grid<false>, lds, L);
            if (st != OHM_OK) return st;
        } else if (need_bwd) {
            OHM_DISPATCH_W(L, (sw_linear_scaled_forward_kernel<true, W>));
        } else {
            OHM_DISPATCH_W(L, (sw_linear_scaled_forward_kernel<false, W>));
        }
        stage_mark(1, L.stream);
        L.a.sel = 1;
        if (need_bwd) OHM_DISPATCH_W(L, (sw_forward_kernel<true, W>));
        else OHM_DISPATCH_W(L, (sw_forward_kernel<false, W>));
    } else {
        if (need_bwd) OHM_DISPATCH_W(L, (sw_forward_kernel<true, W>));
        else OHM_DISPATCH_W(L, (sw_forward_kernel<false, W>));
    }
    if (!scaled && !lin_scaled) stage_mark(1, L.stream);
    stage_mark(2, L.stream);
    if ((st = check_launch("forward kernel")) != OHM_OK) return st;
    if (!need_bwd) return OHM_OK;
    if (scaled) {
        L.a.sel = 0;
        if (grid) {
            OHM_WITH_CK(L.ck, st = launch_grid(sw_affine_scaled_backward_grid<CK>,
                                               grid_bwd_lds_bytes(L.a.max_strips), L));
            if (st != OHM_OK) return st;
        } else if (team) {
            OHM_WITH_CK(L.ck, st = launch_team(sw_affine_scaled_backward_team<CK>,
                                               team_bwd_lds_bytes(L.a.max_strips), L));
            if (st != OHM_OK) return st;
        } else if (!fused) {
            OHM_WITH_CK(L.ck, OHM_DISPATCH_W(L, (sw_affine_scaled_backward_kernel<W, CK>)));
        }
        stage_mark(3, L.stream);
        L.a.sel = 1;
        OHM_DISPATCH_W(L, (sw_affine_backward_kernel<W>));
    } else if (affine) {
        OHM_DISPATCH_W(L, (sw_affine_backward_kernel<W>));
    } else if (lin_scaled) {
        L.a.sel = 0;
        if (lin_grid) {
            st = launch_grid(sw_linear_scaled_backward_grid, grid_bwd_lds_bytes(L.a.max_strips), L);
            if (st != OHM_OK) return st;
        } else {
            OHM_DISPATCH_W(L, (sw_linear_scaled_backward_kernel<W>));
        }
        stage_mark(3, L.stream);
        L.a.sel = 1;
        OHM_DISPATCH_W(L, (sw_backward_kernel<W>));
    } else {
        OHM_DISPATCH_W(L, (sw_backward_kernel<W>));
    }
    if (!scaled && !lin_scaled) stage_mark(3, L.stream);
    stage_mark(4, L.stream);
    return check_launch("backward kernel");
}

// Needleman-Wunsch affine: value, then (when any backward output is asked
// for) map and parameter gradients.
ohm_status nw_forward_impl(const float* scores, const int32_t* lengths, const int64_t* offsets,
                           int B, int L1max, int L2max, float go, float ge, float temp,
                           float* value, float* map, float* g0, float* g1, float* g2, void* ws,
                           size_t ws_bytes, void* stream, bool gap_sum = false) {
    if (B > 0 && !value) return fail(OHM_EINVAL, "value is null");
    const bool need_bwd = map || g0 || g1 || g2;
    if (need_bwd && !map && L1max > 0 && L2max > 0)  // (a batch without cells has no map)
        return fail(OHM_EINVAL, "the gradients need the map buffer");
    Launch L;
    ohm_status st = setup(need_bwd ? kOpNw : kOpValue, true, scores, lengths, offsets, B, L1max, L2max, go,
                          ge, temp, ws, ws_bytes, stream, L);
    if (st != OHM_OK || B == 0) return st;
    L.a.value = value;
    L.a.map = map;
    L.a.g0 = g0;
    L.a.g1 = g1;
    L.a.g2 = g2;
    L.a.gap_sum = gap_sum ? 1 : 0;
    if (need_bwd) OHM_DISPATCH_W(L, (nw_affine_forward_kernel<true, W>));
    else OHM_DISPATCH_W(L, (nw_affine_forward_kernel<false, W>));
    stage_mark(1, L.stream);
    stage_mark(2, L.stream);
    if ((st = check_launch("nw forward kernel")) != OHM_OK || !need_bwd) return st;
    OHM_DISPATCH_W(L, (nw_affine_backward_kernel<W>));
    stage_mark(3, L.stream);
    stage_mark(4, L.stream);
    return check_launch("nw backward kernel");
}

// Needleman-Wunsch affine forward-over-reverse pass along one direction.
ohm_status nw_tangent_impl(const float* scores, const int32_t* lengths, const int64_t* offsets,
                           int B, int L1max, int L2max, float go, float ge, float temp,
                           const float* dir, float dgo, float dge, float dT, float* map_t,
                           float* contract, float* value, float* map, void* ws, size_t ws_bytes,
                           void* stream) {
    if (!std::isfinite(dgo) || !std::isfinite(dge) || !std::isfinite(dT))
        return fail(OHM_EINVAL, "tangent directions must be finite");
    if (contract && (dgo != 0.0f || dge != 0.0f || dT != 0.0f))
        return fail(OHM_EINVAL, "contract needs a pure score direction");
    Launch L;
    ohm_status st = setup(kOpNwTangent, true, scores, lengths, offsets, B, L1max, L2max, go, ge, temp, ws,
                          ws_bytes, stream, L);
    if (st != OHM_OK || B == 0) return st;
    L.a.dir = dir;
    L.a.d_go = dgo;
    L.a.d_ge = dge;
    L.a.d_temp = dT;
    L.a.map_t = map_t;
    L.a.contract = contract;
    L.a.value = value;
    L.a.map = map;
    OHM_DISPATCH_W(L, (nw_affine_tangent_forward_kernel<W>));
    stage_mark(1, L.stream);
    stage_mark(2, L.stream);
    if ((st = check_launch("nw tangent forward")) != OHM_OK) return st;
    OHM_DISPATCH_W(L, (nw_affine_tangent_backward_kernel<W>));
    stage_mark(3, L.stream);
    stage_mark(4, L.stream);
    return check_launch("nw tangent backward");
}

// Saigo-Vert affine forward-over-reverse pass along one direction.
ohm_status sv_tangent_impl(const float* scores, const int32_t* lengths, const int64_t* offsets,
                           int B, int L1max, int L2max, float go, float ge, float temp,
                           const float* dir, float dgo, float dge, float dT, float* map_t,
                           float* contract, float* value, float* map, void* ws, size_t ws_bytes,
                           void* stream) {
    if (!std::isfinite(dgo) || !std::isfinite(dge) || !std::isfinite(dT))
        return fail(OHM_EINVAL, "tangent directions must be finite");
    if (contract && (dgo != 0.0f || dge != 0.0f || dT != 0.0f))
        return fail(OHM_EINVAL, "contract needs a pure score direction");
    Launch L;
    ohm_status st = setup(kOpSvTangent, true, scores, lengths, offsets, B, L1max, L2max, go, ge, temp, ws,
                          ws_bytes, stream, L);
    if (st != OHM_OK || B == 0) return st;
    L.a.dir = dir;
    L.a.d_go = dgo;
    L.a.d_ge = dge;
    L.a.d_temp = dT;
    L.a.map_t = map_t;
    L.a.contract = contract;
    L.a.value = value;
    L.a.map = map;
    OHM_DISPATCH_W(L, (sv_affine_tangent_forward_kernel<W>));
    stage_mark(1, L.stream);
    stage_mark(2, L.stream);
    if ((st = check_launch("sv tangent forward")) != OHM_OK) return st;
    OHM_DISPATCH_W(L, (sv_affine_tangent_backward_kernel<W>));
    stage_mark(3, L.stream);
    stage_mark(4, L.stream);
    return check_launch("sv tangent backward");
}

// Saigo-Vert affine: value, then (when any backward output is asked for) map
// and parameter gradients.
ohm_status sv_forward_impl(const float* scores, const int32_t* lengths, const int64_t* offsets,
                           int B, int L1max, int L2max, float go, float ge, float temp,
                           float* value, float* map, float* g0, float* g1, float* g2, void* ws,
                           size_t ws_bytes, void* stream, bool gap_sum = false) {
    if (B > 0 && !value) return fail(OHM_EINVAL, "value is null");
    const bool need_bwd = map || g0 || g1 || g2;
    if (need_bwd && !map && L1max > 0 && L2max > 0)  // (a batch without cells has no map)
        return fail(OHM_EINVAL, "the gradients need the map buffer");
    Launch L;
    ohm_status st = setup(need_bwd ? kOpSv : kOpValue, true, scores, lengths, offsets, B, L1max, L2max, go,
                          ge, temp, ws, ws_bytes, stream, L);
    if (st != OHM_OK || B == 0) return st;
    L.a.value = value;
    L.a.map = map;
    L.a.g0 = g0;
    L.a.g1 = g1;
    L.a.g2 = g2;
    L.a.gap_sum = gap_sum ? 1 : 0;
    if (need_bwd) OHM_DISPATCH_W(L, (sv_affine_forward_kernel<true, W>));
    else OHM_DISPATCH_W(L, (sv_affine_forward_kernel<false, W>));
    stage_mark(1, L.stream);
    stage_mark(2, L.stream);
    if ((st = check_launch("sv forward kernel")) != OHM_OK || !need_bwd) return st;
    OHM_DISPATCH_W(L, (sv_affine_backward_kernel<W>));
    stage_mark(3, L.stream);
    stage_mark(4, L.stream);
    return check_launch("sv backward kernel");
}

ohm_status tangent_impl(bool affine, const float* scores, const int32_t* lengths,
                        const int64_t* offsets, int B, int L1max, int L2max, float go, float ge,
                        float temp, const float* dir, float dgo, float dge, float dT,
                        float* map_t, float* contract, float* value, float* map, void* ws,
                        size_t ws_bytes, void* stream) {
    if (!std::isfinite(dgo) || !std::isfinite(dge) || !std::isfinite(dT))
        return fail(OHM_EINVAL, "tangent directions must be finite");
    if (contract && (dgo != 0.0f || dge != 0.0f || dT != 0.0f))
        return fail(OHM_EINVAL, "contract needs a pure score direction");
    Launch L;
    ohm_status st = setup(kOpTangent, affine, scores, lengths, offsets, B, L1max, L2max, go, ge, temp, ws,
                          ws_bytes, stream, L);
    if (st != OHM_OK || B == 0) return st;
    L.a.dir = dir;
    L.a.d_go = dgo;
    L.a.d_ge = dge;
    L.a.d_temp = dT;
    L.a.map_t = map_t;
    L.a.contract = contract;
    L.a.value = value;
    L.a.map = map;
    if (affine && use_scaled(L.a)) {
        // exponent-tracked tangent sweeps for every pair, then the log-domain
        // tangent sweeps for the pairs they flagged
        const int mode = contract ? kTanContract
                                  : (dgo != 0.0f || dge != 0.0f || dT != 0.0f) ? kTanParam : kTanScore;
        const bool grid = use_grid(L.a);
        const size_t lds = grid_cum_bytes(L.a.max_strips);
        if (grid && (st = launch_grid_fill(L)) != OHM_OK) return st;
        if (mode == kTanParam) {
            if (grid) st = launch_grid(sw_affine_param_grad_scaled_forward_grid, lds, L);
            else OHM_DISPATCH_W(L, (sw_affine_param_grad_scaled_forward_kernel<W>));
        } else {
            if (grid) st = launch_grid(sw_affine_hvp_scaled_forward_grid, lds, L);
            else OHM_DISPATCH_W(L, (sw_affine_hvp_scaled_forward_kernel<W>));
        }
        if (st != OHM_OK) return st;
        stage_mark(1, L.stream);
        L.a.sel = 1;
        OHM_DISPATCH_W(L, (sw_affine_hvp_forward_kernel<W>));
        stage_mark(2, L.stream);
        if ((st = check_launch("tangent forward")) != OHM_OK) return st;
        L.a.sel = 0;
        if (grid && (st = launch_grid_fill(L)) != OHM_OK) return st;
        if (mode == kTanParam) {
            if (grid) st = launch_grid(sw_affine_param_grad_scaled_backward_grid, lds, L);
            else OHM_DISPATCH_W(L, (sw_affine_param_grad_scaled_backward_kernel<W>));
        } else if (mode == kTanContract) {
            if (grid) st = launch_grid(sw_affine_param_grad_contract_backward_grid, lds, L);
            else OHM_DISPATCH_W(L, (sw_affine_param_grad_contract_backward_kernel<W>));
        } else {
            if (grid) st = launch_grid(sw_affine_hvp_scaled_backward_grid, lds, L);
            else OHM_DISPATCH_W(L, (sw_affine_hvp_scaled_backward_kernel<W>));
        }
        if (st != OHM_OK) return st;
        stage_mark(3, L.stream);
        L.a.sel = 1;
        OHM_DISPATCH_W(L, (sw_affine_hvp_backward_kernel<W>));
        stage_mark(4, L.stream);
        return check_launch("tangent backward");
    } else if (affine && use_grid(L.a)) {
        // the tangent sweeps' entries share words 0..5: a fill before each
        const size_t lds = grid_cum_bytes(L.a.max_strips);
        if ((st = launch_grid_fill(L)) != OHM_OK) return st;
        if ((st = launch_grid(sw_affine_tangent_forward_grid, lds, L)) != OHM_OK) return st;
        stage_mark(1, L.stream);
        stage_mark(2, L.stream);
        if ((st = launch_grid_fill(L)) != OHM_OK) return st;
        if ((st = launch_grid(sw_affine_tangent_backward_grid, lds, L)) != OHM_OK) return st;
    } else if (affine) {
        if (dir) {
            OHM_DISPATCH_W(L, (sw_affine_hvp_forward_kernel<W>));
            stage_mark(1, L.stream);
            stage_mark(2, L.stream);
            if ((st = check_launch("tangent forward")) != OHM_OK) return st;
            OHM_DISPATCH_W(L, (sw_affine_hvp_backward_kernel<W>));
        } else {
            OHM_DISPATCH_W(L, (sw_affine_param_grad_forward_kernel<W>));
            stage_mark(1, L.stream);
            stage_mark(2, L.stream);
            if ((st = check_launch("tangent forward")) != OHM_OK) return st;
            OHM_DISPATCH_W(L, (sw_affine_param_grad_backward_kernel<W>));
        }
    } else {
        OHM_DISPATCH_W(L, (sw_hvp_forward_kernel<W>));
        stage_mark(1, L.stream);
        stage_mark(2, L.stream);
        if ((st = check_launch("tangent forward")) != OHM_OK) return st;
        OHM_DISPATCH_W(L, (sw_hvp_backward_kernel<W>));
    }
    stage_mark(3, L.stream);
    stage_mark(4, L.stream);
    return check_launch("tangent backward");
}

}  // namespace

extern "C" ohm_status ohm_set_stage_events(void* const* events, int32_t count) {
    if (count < 0 || count > kStageMarks || (count > 0 && !events))
        return fail(OHM_EINVAL, "stage events: count must be 0..5 with a non-null array");
    for (int k = 0; k < kStageMarks; ++k) g_stage_events[k] = k < count ? events[k] : nullptr;
    return OHM_OK;
}

extern "C" ohm_status ohm_set_launch_config(int32_t waves, int32_t groups_per_cu, int32_t teams) {
    if (!(waves == 0 || waves == 1 || waves == 2 || waves == 4 || waves == 8))
        return fail(OHM_EINVAL, "launch config: waves must be 0, 1, 2, 4 or 8");
    if (groups_per_cu < 0 || groups_per_cu > 64)
        return fail(OHM_EINVAL, "launch config: groups_per_cu must be 0..64");
    if (teams < 0 || teams > 4) return fail(OHM_EINVAL, "launch config: teams must be 0..4");
    g_cfg_waves.store(waves);
    g_cfg_per_cu.store(groups_per_cu);
    g_cfg_teams.store(teams);
    return OHM_OK;
}

extern "C" ohm_status ohm_set_checkpoint_stride(int32_t steps) {
    if (!(steps == 0 || steps == 4 || steps == 8 || steps == 16))
        return fail(OHM_EINVAL, "checkpoint stride must be 0 (automatic), 4, 8 or 16");
    g_cfg_ckpt.store(steps);
    return OHM_OK;
}

extern "C" int32_t ohm_checkpoint_stride(void) { return ckpt_stride(); }

extern "C" size_t ohm_sw_affine_workspace(int32_t op, int32_t B, int32_t L1max, int32_t L2max,
                                          int64_t total_cells) {
    if (B < 0 || L1max < 0 || L2max < 0 || op < 0 || op > 2) return 0;
    return ws_layout(op, B, L1max, L2max, total_cells,
                     TableSpec{kPlanesAffineCkpt, affine_table_stride(ckpt_stride())},
                     TableSpec{kPlanesAffineTan, 1}).total;
}

extern "C" size_t ohm_nw_affine_workspace(int32_t op, int32_t B, int32_t L1max, int32_t L2max,
                                          int64_t total_cells) {
    if (B < 0 || L1max < 0 || L2max < 0 || op < 0 || op > 2) return 0;
    return ws_layout(op, B, L1max, L2max, total_cells, TableSpec{kPlanesNw, 1},
                     TableSpec{kPlanesNwTan, 1}).total;
}

extern "C" ohm_status ohm_nw_affine_tangent(const float* scores, const int32_t* lengths,
                                            const int64_t* offsets, int32_t B, int32_t L1max,
                                            int32_t L2max, float gap_open, float gap_ext,
                                            float temp, const float* direction, float d_gap_open,
                                            float d_gap_ext, float d_temp, float* map_tangent,
                                            float* contract, float* value, float* map,
                                            void* workspace, size_t workspace_bytes,
                                            void* stream) {
    ohm_status st = nw_tangent_impl(scores, lengths, offsets, B, L1max, L2max, gap_open, gap_ext, temp,
                           direction, d_gap_open, d_gap_ext, d_temp, map_tangent, contract,
                           value, map, workspace, workspace_bytes, stream);
    if (st != OHM_OK) return st;
    return finish_lost(lengths, offsets, B, L1max, L2max, workspace, CallOut{value, {nullptr, nullptr, nullptr}, contract, 3, map, map_tangent}, stream);
}

extern "C" ohm_status ohm_nw_affine_forward(const float* scores, const int32_t* lengths,
                                            const int64_t* offsets, int32_t B, int32_t L1max,
                                            int32_t L2max, float gap_open, float gap_ext,
                                            float temp, float* value, float* map,
                                            float* grad_open, float* grad_ext, float* grad_temp,
                                            void* workspace, size_t workspace_bytes,
                                            void* stream) {
    ohm_status st = nw_forward_impl(scores, lengths, offsets, B, L1max, L2max, gap_open, gap_ext, temp,
                           value, map, grad_open, grad_ext, grad_temp, workspace,
                           workspace_bytes, stream);
    if (st != OHM_OK) return st;
    return finish_lost(lengths, offsets, B, L1max, L2max, workspace, CallOut{value, {grad_open, grad_ext, grad_temp}, nullptr, 0, map, nullptr}, stream);
}

extern "C" size_t ohm_sv_affine_workspace(int32_t op, int32_t B, int32_t L1max, int32_t L2max,
                                          int64_t total_cells) {
    if (B < 0 || L1max < 0 || L2max < 0 || op < 0 || op > 2) return 0;
    return ws_layout(op, B, L1max, L2max, total_cells, TableSpec{kPlanesSv, 1},
                     TableSpec{kPlanesSvTan, 1}).total;
}

extern "C" ohm_status ohm_sv_affine_tangent(const float* scores, const int32_t* lengths,
                                            const int64_t* offsets, int32_t B, int32_t L1max,
                                            int32_t L2max, float gap_open, float gap_ext,
                                            float temp, const float* direction, float d_gap_open,
                                            float d_gap_ext, float d_temp, float* map_tangent,
                                            float* contract, float* value, float* map,
                                            void* workspace, size_t workspace_bytes,
                                            void* stream) {
    ohm_status st = sv_tangent_impl(scores, lengths, offsets, B, L1max, L2max, gap_open, gap_ext, temp,
                           direction, d_gap_open, d_gap_ext, d_temp, map_tangent, contract,
                           value, map, workspace, workspace_bytes, stream);
    if (st != OHM_OK) return st;
    return finish_lost(lengths, offsets, B, L1max, L2max, workspace, CallOut{value, {nullptr, nullptr, nullptr}, contract, 3, map, map_tangent}, stream);
}

extern "C" ohm_status ohm_sv_affine_forward(const float* scores, const int32_t* lengths,
                                            const int64_t* offsets, int32_t B, int32_t L1max,
                                            int32_t L2max, float gap_open, float gap_ext,
                                            float temp, float* value, float* map,
                                            float* grad_open, float* grad_ext, float* grad_temp,
                                            void* workspace, size_t workspace_bytes,
                                            void* stream) {
    ohm_status st = sv_forward_impl(scores, lengths, offsets, B, L1max, L2max, gap_open, gap_ext, temp,
                           value, map, grad_open, grad_ext, grad_temp, workspace,
                           workspace_bytes, stream);
    if (st != OHM_OK) return st;
    return finish_lost(lengths, offsets, B, L1max, L2max, workspace, CallOut{value, {grad_open, grad_ext, grad_temp}, nullptr, 0, map, nullptr}, stream);
}

// Linear-gap Needleman-Wunsch and Saigo-Vert: the affine sweeps with one gap
// score in both roles (every move sequence has one state sequence, so the two
// models are the same function); dS/dgap = dS/dgap_open + dS/dgap_ext.
extern "C" size_t ohm_nw_workspace(int32_t op, int32_t B, int32_t L1max, int32_t L2max,
                                   int64_t total_cells) {
    return ohm_nw_affine_workspace(op, B, L1max, L2max, total_cells);
}

extern "C" size_t ohm_sv_workspace(int32_t op, int32_t B, int32_t L1max, int32_t L2max,
                                   int64_t total_cells) {
    return ohm_sv_affine_workspace(op, B, L1max, L2max, total_cells);
}

extern "C" ohm_status ohm_nw_forward(const float* scores, const int32_t* lengths,
                                     const int64_t* offsets, int32_t B, int32_t L1max,
                                     int32_t L2max, float gap, float temp, float* value,
                                     float* map, float* grad_gap, float* grad_temp,
                                     void* workspace, size_t workspace_bytes, void* stream) {
    ohm_status st = nw_forward_impl(scores, lengths, offsets, B, L1max, L2max, gap, gap, temp, value, map,
                           grad_gap, nullptr, grad_temp, workspace, workspace_bytes, stream, true);
    if (st != OHM_OK) return st;
    return finish_lost(lengths, offsets, B, L1max, L2max, workspace, CallOut{value, {grad_gap, grad_temp, nullptr}, nullptr, 0, map, nullptr}, stream);
}

extern "C" ohm_status ohm_sv_forward(const float* scores, const int32_t* lengths,
                                     const int64_t* offsets, int32_t B, int32_t L1max,
                                     int32_t L2max, float gap, float temp, float* value,
                                     float* map, float* grad_gap, float* grad_temp,
                                     void* workspace, size_t workspace_bytes, void* stream) {
    ohm_status st = sv_forward_impl(scores, lengths, offsets, B, L1max, L2max, gap, gap, temp, value, map,
                           grad_gap, nullptr, grad_temp, workspace, workspace_bytes, stream, true);
    if (st != OHM_OK) return st;
    return finish_lost(lengths, offsets, B, L1max, L2max, workspace, CallOut{value, {grad_gap, grad_temp, nullptr}, nullptr, 0, map, nullptr}, stream);
}

extern "C" size_t ohm_sw_workspace(int32_t op, int32_t B, int32_t L1max, int32_t L2max,
                                   int64_t total_cells) {
    if (B < 0 || L1max < 0 || L2max < 0 || op < 0 || op > 2) return 0;
    return ws_layout(op, B, L1max, L2max, total_cells, TableSpec{kPlanesLinearFwd, 1},
                     TableSpec{kPlanesLinearTan, 1}).total;
}

extern "C" ohm_status ohm_sw_affine_forward(const float* scores, const int32_t* lengths,
                                            const int64_t* offsets, int32_t B, int32_t L1max,
                                            int32_t L2max, float gap_open, float gap_ext,
                                            float temp, float* value, float* map,
                                            float* grad_open, float* grad_ext, float* grad_temp,
                                            void* workspace, size_t workspace_bytes,
                                            void* stream) {
    ohm_status st = forward_impl(true, scores, lengths, offsets, B, L1max, L2max, gap_open, gap_ext, temp,
                        value, map, grad_open, grad_ext, grad_temp, workspace, workspace_bytes,
                        stream);
    if (st != OHM_OK) return st;
    return finish_lost(lengths, offsets, B, L1max, L2max, workspace, CallOut{value, {grad_open, grad_ext, grad_temp}, nullptr, 0, map, nullptr}, stream);
}

extern "C" ohm_status ohm_sw_affine_tangent(const float* scores, const int32_t* lengths,
                                            const int64_t* offsets, int32_t B, int32_t L1max,
                                            int32_t L2max, float gap_open, float gap_ext,
                                            float temp, const float* direction, float d_gap_open,
                                            float d_gap_ext, float d_temp, float* map_tangent,
                                            float* contract, float* value, float* map,
                                            void* workspace, size_t workspace_bytes,
                                            void* stream) {
    ohm_status st = tangent_impl(true, scores, lengths, offsets, B, L1max, L2max, gap_open, gap_ext, temp,
                        direction, d_gap_open, d_gap_ext, d_temp, map_tangent, contract, value,
                        map, workspace, workspace_bytes, stream);
    if (st != OHM_OK) return st;
    return finish_lost(lengths, offsets, B, L1max, L2max, workspace, CallOut{value, {nullptr, nullptr, nullptr}, contract, 3, map, map_tangent}, stream);
}

extern "C" ohm_status ohm_sw_forward(const float* scores, const int32_t* lengths,
                                     const int64_t* offsets, int32_t B, int32_t L1max,
                                     int32_t L2max, float gap, float temp, float* value,
                                     float* map, float* grad_gap, float* grad_temp,
                                     void* workspace, size_t workspace_bytes, void* stream) {
    ohm_status st = forward_impl(false, scores, lengths, offsets, B, L1max, L2max, gap, 0.0f, temp, value,
                        map, grad_gap, grad_temp, nullptr, workspace, workspace_bytes, stream);
    if (st != OHM_OK) return st;
    return finish_lost(lengths, offsets, B, L1max, L2max, workspace, CallOut{value, {grad_gap, grad_temp, nullptr}, nullptr, 0, map, nullptr}, stream);
}

extern "C" ohm_status ohm_sw_tangent(const float* scores, const int32_t* lengths,
                                     const int64_t* offsets, int32_t B, int32_t L1max,
                                     int32_t L2max, float gap, float temp, const float* direction,
                                     float d_gap, float d_temp, float* map_tangent,
                                     float* contract, float* value, float* map, void* workspace,
                                     size_t workspace_bytes, void* stream) {
    ohm_status st = tangent_impl(false, scores, lengths, offsets, B, L1max, L2max, gap, 0.0f, temp,
                        direction, d_gap, 0.0f, d_temp, map_tangent, contract, value, map,
                        workspace, workspace_bytes, stream);
    if (st != OHM_OK) return st;
    return finish_lost(lengths, offsets, B, L1max, L2max, workspace, CallOut{value, {nullptr, nullptr, nullptr}, contract, 2, map, map_tangent}, stream);
}

// --------------------------------------------------------------------------
// Embedding-scored affine sweeps, scores generated inside the sweep
// (include/ohm_embed.h; kernels in embed_fused.hiph)
// --------------------------------------------------------------------------

extern "C" ohm_status ohm_embed_offsets(const int32_t* lengths, int32_t B, int32_t Lq, int32_t Lk,
                                        int64_t* offsets, void* stream);

extern "C" ohm_status ohm_embed_sw_affine_forward(const void* Q, const void* K, const int32_t* lengths,
                                                  int32_t B, int32_t Lq, int32_t Lk, int32_t D,
                                                  float gap_open, float gap_ext, float temp,
                                                  int64_t* offsets, float* scores, float* value,
                                                  float* map, float* grad_open, float* grad_ext,
                                                  float* grad_temp, void* workspace,
                                                  size_t workspace_bytes, void* stream) {
    if (B < 0 || Lq < 0 || Lk < 0) return fail(OHM_EINVAL, "negative batch size or length");
    if (!(D == 32 || D == 64 || D == 128))
        return fail(OHM_EINVAL, "the fused embedding sweeps take D = 32, 64 or 128 "
                                "(D = 256: ohm_embed_scores, then ohm_sw_affine_forward)");
    if (B == 0) return OHM_OK;
    if (!lengths || !offsets || !value) return fail(OHM_EINVAL, "lengths, offsets and value must be non-null");
    if (Lq > 0 && Lk > 0 && (!Q || !K || !scores)) return fail(OHM_EINVAL, "Q, K and scores must be non-null");
    ohm_status st = ohm_embed_offsets(lengths, B, Lq, Lk, offsets, stream);
    if (st != OHM_OK) return st;
    Launch L;
    st = setup(kOpForward, true, scores, lengths, offsets, B, Lq, Lk, gap_open, gap_ext, temp, workspace,
               workspace_bytes, stream, L);
    if (st != OHM_OK) return st;
    if (!use_scaled(L.a))
        return fail(OHM_EINVAL, "gap factors outside the exponent-tracked sweeps' range "
                                "(ohm_embed_scores, then ohm_sw_affine_forward)");
    L.a.value = value;
    L.a.map = map;
    L.a.g0 = grad_open;
    L.a.g1 = grad_ext;
    L.a.g2 = grad_temp;
    const EmbArgs e{static_cast<const __bf16*>(Q), static_cast<const __bf16*>(K), scores, Lq, Lk};
    constexpr int kW = 4;
    const size_t lds = embed_fused_lds_bytes(kW, L.a.max_strips);
    // two workgroups per CU fit the LDS (rings and map rings ~67 KB each)
    const unsigned groups = (unsigned)std::max(1, std::min(B, 2 * device_cus()));
    auto launch = [&](auto kernel) {
        if (hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
            return fail(OHM_ELAUNCH, "cannot reserve LDS for the embedding sweep kernel");
        hipLaunchKernelGGL(kernel, dim3(groups), dim3(64 * kW), lds, L.stream, L.a, e);
        return check_launch("embedding sweep kernel");
    };
    if (D == 32) st = launch(embed_sw_fused_kernel<kW, 32>);
    else if (D == 64) st = launch(embed_sw_fused_kernel<kW, 64>);
    else st = launch(embed_sw_fused_kernel<kW, 128>);
    if (st != OHM_OK) return st;
    stage_mark(1, L.stream);
    stage_mark(2, L.stream);
    // pairs the scaled sweep flagged: the log-domain sweeps, on the scores
    // the fused kernel wrote for them
    L.a.sel = 1;
    OHM_DISPATCH_W(L, (sw_affine_forward_kernel<true, W>));
    if ((st = check_launch("embedding fallback forward")) != OHM_OK) return st;
    stage_mark(3, L.stream);
    OHM_DISPATCH_W(L, (sw_affine_backward_kernel<W>));
    stage_mark(4, L.stream);
    if ((st = check_launch("embedding fallback backward")) != OHM_OK) return st;
    return finish_lost(lengths, offsets, B, Lq, Lk, workspace,
                       CallOut{value, {grad_open, grad_ext, grad_temp}, nullptr, 0, map, nullptr}, stream);
}

// --------------------------------------------------------------------------
// Allocation-free variant of ohm_sw_affine_forward: the workspace is a
// fixed device array of this library, so callers that cannot allocate (code
// under HIP graph capture, serving loops that forbid hipMalloc) can run the
// sweeps.  Each GPU the library runs on holds its own copy of the array for
// the life of the process; calls on one GPU share it, so calls that overlap
// on one GPU race.  Every byte the kernels read from it is written first by
// the same call (the planning kernel lays out and clears the claims, then
// the sweeps write each table and boundary block before reading it).
// --------------------------------------------------------------------------

constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kScratchDevices];  // device address of `scratch`, per device

// The array's device address on the current device (looked up once per
// device; the host-side symbol is only a handle, never a device address).
ohm_status current_scratch(void** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(OHM_ELAUNCH, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kScratchDevices) return fail(OHM_ELAUNCH, "device index out of range for the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(OHM_ELAUNCH, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return OHM_OK;
}

}  // namespace

extern "C" size_t ohm_scratch_bytes(void) { return kScratchBytes; }

extern "C" ohm_status ohm_sw_affine_forward_noalloc(const float* scores, const int32_t* lengths,
                                                    const int64_t* offsets, int64_t total_cells,
                                                    int32_t B, int32_t L1max, int32_t L2max,
                                                    float gap_open, float gap_ext, float temp,
                                                    float* value, float* map, float* grad_open,
                                                    float* grad_ext, float* grad_temp, void* stream) {
    if (offsets && total_cells < 0)
        return fail(OHM_EINVAL, "the packed layout needs total_cells (the sum of L1_b * L2_b)");
    const int op = (map || grad_open || grad_ext || grad_temp) ? OHM_OP_FORWARD : OHM_OP_VALUE;
    const size_t need = ohm_sw_affine_workspace(op, B, L1max, L2max, offsets ? total_cells : -1);
    if (need > kScratchBytes)
        return fail(OHM_EWORKSPACE, "batch needs " + std::to_string(need) + " bytes of workspace; the "
                    "allocation-free entry point has " + std::to_string(kScratchBytes) +
                    " (use ohm_sw_affine_forward with a caller workspace)");
    void* ws = nullptr;
    ohm_status st = current_scratch(&ws);
    if (st != OHM_OK) return st;
    return ohm_sw_affine_forward(scores, lengths, offsets, B, L1max, L2max, gap_open, gap_ext, temp,
                                 value, map, grad_open, grad_ext, grad_temp, ws, kScratchBytes,
                                 stream);
}
