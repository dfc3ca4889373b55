// Embedding-scored alignment on the gfx950 matrix cores (include/ohm_embed.h).
//
// Three ragged BF16 MFMA kernels, each touching only a pair's own
// M_b x N_b box:
//   scores  s_b = Q_b K_b^T                 written packed (the sweeps' layout)
//   dQ      dQ_b^T = K_b^T (g_b P_b)^T       reduction over the pair's columns
//   dK      dK_b^T = Q_b^T (g_b P_b)         reduction over the pair's rows
// All three use v_mfma_f32_16x16x32_bf16.  Operand layout (lane l, r = l%16,
// q = l/16): A holds row r, reduction slots 8q..8q+7; B holds column r, the
// same slots; the 16x16 FP32 result holds rows 4q..4q+3 of column r.  The
// reduction slots of A and B are paired by the instruction, so any fixed
// assignment of reduction indices to (q, slot) is a valid product.
//
// Each wave loads its operands straight from global memory: a reduction
// slot run is 16 contiguous bytes (one dwordx4) when the reduction index is
// the fast axis of the array.  For dQ and dK that axis is the embedding
// dimension of Q and K, so the kernels read transposed copies (Q^T, K^T,
// zero past each pair's length) written by embed_transpose_kernel.  The
// map P is FP32; it enters the BF16 products as two terms, P = hi + lo with
// hi = bf16(P) and lo = bf16(P - hi), which keeps ~16 bits of it.
//
// Work is dealt to the 8 XCDs in contiguous runs (work item w runs on XCD
// w / (grid / 8)), so the tiles of one pair share an L2 and its Q and K rows
// are fetched from HBM about once.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "ohm_embed.h"

namespace ohm {
void set_error(const std::string& msg);
}

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

constexpr int kRows = 128;     // output rows per workgroup (32 per wave)
constexpr int kThreads = 256;  // four waves
constexpr int kXcds = 8;

#define DEV static __device__ __forceinline__

DEV int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

DEV f32x4 mma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// FP32 -> BF16 bits, round to nearest even (finite inputs)
DEV uint32_t bf16_bits(float x) {
    uint32_t u = __builtin_bit_cast(uint32_t, x);
    return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}
DEV float bf16_value(uint32_t h) { return __builtin_bit_cast(float, h << 16); }

// p = hi + lo as two BF16 vectors
DEV void split8(const float (&p)[8], bf16x8& hi, bf16x8& lo) {
    uint32_t h[8], l[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        h[v] = bf16_bits(p[v]);
        l[v] = bf16_bits(p[v] - bf16_value(h[v]));
    }
    u32x4 hv, lv;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        hv[v] = h[2 * v] | (h[2 * v + 1] << 16);
        lv[v] = l[2 * v] | (l[2 * v + 1] << 16);
    }
    hi = __builtin_bit_cast(bf16x8, hv);
    lo = __builtin_bit_cast(bf16x8, lv);
}

// four FP32 values as four BF16 (8 bytes)
DEV u32x2 pack4(float a, float b, float c, float d) {
    u32x2 r;
    r[0] = bf16_bits(a) | (bf16_bits(b) << 16);
    r[1] = bf16_bits(c) | (bf16_bits(d) << 16);
    return r;
}

// work item of this workgroup: contiguous runs of items per XCD (the
// hardware deals workgroup ids to the XCDs round robin); grid % 8 == 0
DEV int work_item() {
    const int per = gridDim.x / kXcds;
    return (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
}

int round_grid(int64_t items) { return (int)((items + kXcds - 1) / kXcds * kXcds); }

// --------------------------------------------------------------------------
// offsets[b] = sum_{b' < b} M_b' N_b' (one workgroup; exact in int64)
// --------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void embed_offsets_kernel(const int* __restrict__ lengths, int B,
                                                             int Lq, int Lk, int64_t* __restrict__ offsets) {
    __shared__ int64_t part[1024];
    const int t = threadIdx.x, n = blockDim.x;
    const int per = (B + n - 1) / n;
    const int lo = min(B, t * per), hi = min(B, lo + per);
    int64_t s = 0;
    for (int b = lo; b < hi; ++b)
        s += (int64_t)clampi(lengths[2 * b], 0, Lq) * clampi(lengths[2 * b + 1], 0, Lk);
    part[t] = s;
    __syncthreads();
    for (int d = 1; d < n; d <<= 1) {  // inclusive scan
        const int64_t v = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int64_t run = t > 0 ? part[t - 1] : 0;
    for (int b = lo; b < hi; ++b) {
        offsets[b] = run;
        run += (int64_t)clampi(lengths[2 * b], 0, Lq) * clampi(lengths[2 * b + 1], 0, Lk);
    }
}

// --------------------------------------------------------------------------
// scores: one workgroup per 128 rows of a pair, a wave per 32 rows; the wave
// keeps its rows of Q in registers and walks the pair's columns 32 at a
// time, requesting the next 32 rows of K while it multiplies the current
// --------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(kThreads) void embed_scores_kernel(
    const __bf16* __restrict__ Q, const __bf16* __restrict__ K, const int* __restrict__ lengths,
    const int64_t* __restrict__ offsets, int Lq, int Lk, int tiles_i, int items,
    float* __restrict__ S) {
    constexpr int KS = D / 32;  // reduction steps
    const int w = work_item();
    if (w >= items) return;
    const int b = w / tiles_i, t = w - b * tiles_i;
    const int M = clampi(lengths[2 * b], 0, Lq), N = clampi(lengths[2 * b + 1], 0, Lk);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
    const int i0 = t * (4 * 32) + 32 * wave;
    if (i0 >= M || N == 0) return;  // no barrier in this kernel
    bf16x8 a[2][KS];
#pragma unroll
    for (int x = 0; x < 2; ++x) {
        const __bf16* qr = Q + ((size_t)b * Lq + min(i0 + 16 * x + r, M - 1)) * D + 8 * q;
#pragma unroll
        for (int k = 0; k < KS; ++k) a[x][k] = *reinterpret_cast<const bf16x8*>(qr + 32 * k);
    }
    const __bf16* Kb = K + (size_t)b * Lk * D + 8 * q;
    auto load_k = [&](bf16x8 (&dst)[2][KS], int j0) {
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const __bf16* kr = Kb + (size_t)min(j0 + 16 * y + r, N - 1) * D;
#pragma unroll
            for (int k = 0; k < KS; ++k) dst[y][k] = *reinterpret_cast<const bf16x8*>(kr + 32 * k);
        }
    };
    float* Sb = S + offsets[b];
    bf16x8 bc[2][KS], bn[2][KS];
    load_k(bc, 0);
    for (int j0 = 0; j0 < N; j0 += 32) {
        if (j0 + 32 < N) load_k(bn, j0 + 32);
        f32x4 acc[2][2];
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y) acc[x][y] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int k = 0; k < KS; ++k)
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y) acc[x][y] = mma(a[x][k], bc[y][k], acc[x][y]);
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const int col = j0 + 16 * y + r;
            if (col >= N) continue;
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int row = i0 + 16 * x + 4 * q + v;
                    if (row < M) Sb[(size_t)row * N + col] = acc[x][y][v];
                }
        }
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int k = 0; k < KS; ++k) bc[y][k] = bn[y][k];
    }
}

// --------------------------------------------------------------------------
// X [B, L, D] -> XT [B, D, L32] (L32 = L rounded up to 32, so a wave's
// 32-wide reduction run never leaves a row), zero for rows past each pair's
// length (lengths[2 b + side]) and in the padding
// --------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(kThreads) void embed_transpose_kernel(const __bf16* __restrict__ X,
                                                                   const int* __restrict__ lengths,
                                                                   int side, int L, int L32,
                                                                   __bf16* __restrict__ XT) {
    const int b = blockIdx.y;
    const int len = clampi(lengths[2 * b + side], 0, L);
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    if (j >= L32) return;
    const __bf16* src = X + ((size_t)b * L + min(j, L - 1)) * D;
    __bf16* dst = XT + (size_t)b * D * L32 + j;
    for (int c = threadIdx.x >> 6; c < D / 8; c += kThreads / 64) {
        u32x4 v = u32x4{0u, 0u, 0u, 0u};
        if (j < len) v = *reinterpret_cast<const u32x4*>(src + 8 * c);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const uint16_t h = (uint16_t)(v[e >> 1] >> (16 * (e & 1)));
            dst[(size_t)(8 * c + e) * L32] = __builtin_bit_cast(__bf16, h);
        }
    }
}

// --------------------------------------------------------------------------
// dQ and dK: one workgroup per 128 output rows of a pair, a wave per 32 rows
// (two 16-row blocks) x all D.
//   dQ^T[d, i] = sum_j K^T[d, j] P[i, j]   (map rows: 8 contiguous floats)
//   dK^T[d, j] = sum_i Q^T[d, i] P[i, j]   (map columns: 8 rows, 64-byte
//                                           segments across 16 lanes)
// Per reduction step of 32, the D x 32 tile of the transposed copy is the
// same for the four waves: the workgroup loads it once (32 bytes a thread)
// into one of two LDS buffers (rows padded to 80 bytes, so the 16 rows a
// wave reads at once fall in different bank groups), one step ahead; each
// wave keeps its own map values in registers, also one step ahead.  One
// barrier per step separates the buffer's last readers from its next writer.
// --------------------------------------------------------------------------
constexpr int kTileRowBytes = 80;  // 64 bytes of a 32-wide BF16 run + padding

template <int D, bool ROWS>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(D <= 128 ? 3 : 2))) void embed_grad_kernel(
    const __bf16* __restrict__ XT, const float* __restrict__ P, const int* __restrict__ lengths,
    const int64_t* __restrict__ offsets, const float* __restrict__ gv, int Lq, int Lk, int L32,
    int tiles, int items, __bf16* __restrict__ out) {
    constexpr int NB = D / 16;
    constexpr int kChunks = D * 4;  // 16-byte pieces of one D x 32 tile
    constexpr int kPer = (kChunks + kThreads - 1) / kThreads;
    __shared__ __align__(16) unsigned char tile[2][D * kTileRowBytes];
    const int w = work_item();
    if (w >= items) return;  // uniform over the workgroup
    const int b = w / tiles, t = w - b * tiles;
    const int M = clampi(lengths[2 * b], 0, Lq), N = clampi(lengths[2 * b + 1], 0, Lk);
    // ROWS: the output rows are the map's rows (dQ), else its columns (dK)
    const int Lout = ROWS ? Lq : Lk, nout = ROWS ? M : N, nred = ROWS ? N : M;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
    const int o0 = t * (4 * 32) + 32 * wave;
    f32x4 acc[2][NB];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int d = 0; d < NB; ++d) acc[x][d] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    // the workgroup runs the reduction when any of its rows is inside the pair
    if (t * (4 * 32) < nout && nred > 0) {
        const float* Pb = P + offsets[b];
        const bool mine = o0 < nout;  // this wave has rows inside the pair
        bool live[2];
        const float* pp[2];
#pragma unroll
        for (int x = 0; x < 2; ++x) {
            const int o = o0 + 16 * x + r;
            live[x] = o < nout;
            // dQ: row o of the map; dK: column o
            pp[x] = ROWS ? Pb + (size_t)min(o, nout - 1) * N : Pb + min(o, nout - 1);
        }
        auto load_p = [&](int s0, float (&pv)[2][8]) {
            const int ss = s0 + 8 * q;
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                if constexpr (ROWS) {
                    if (live[x] && ss + 8 <= nred) {
                        const f4u u0 = *reinterpret_cast<const f4u*>(pp[x] + ss);
                        const f4u u1 = *reinterpret_cast<const f4u*>(pp[x] + ss + 4);
#pragma unroll
                        for (int v = 0; v < 4; ++v) { pv[x][v] = u0[v]; pv[x][4 + v] = u1[v]; }
                    } else {
#pragma unroll
                        for (int v = 0; v < 8; ++v) pv[x][v] = (live[x] && ss + v < nred) ? pp[x][ss + v] : 0.0f;
                    }
                } else {
#pragma unroll
                    for (int v = 0; v < 8; ++v)
                        pv[x][v] = (live[x] && ss + v < nred) ? pp[x][(size_t)(ss + v) * N] : 0.0f;
                }
            }
        };
        // the workgroup's share of one D x 32 tile: piece c = thread + 256 k is
        // row c / 4, bytes 16 (c % 4) of the row's 64
        const __bf16* xt = XT + (size_t)b * D * L32;
        auto load_tile = [&](int s0, u32x4 (&st)[kPer]) {
#pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const int c = threadIdx.x + kThreads * k;
                if (c < kChunks)
                    st[k] = *reinterpret_cast<const u32x4*>(xt + (size_t)(c >> 2) * L32 + s0 + 8 * (c & 3));
            }
        };
        auto store_tile = [&](int buf, const u32x4 (&st)[kPer]) {
#pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const int c = threadIdx.x + kThreads * k;
                if (c < kChunks)
                    *reinterpret_cast<u32x4*>(&tile[buf][(c >> 2) * kTileRowBytes + 16 * (c & 3)]) = st[k];
            }
        };
        u32x4 stage[kPer];
        float pc[2][8], pn[2][8];
        load_tile(0, stage);
        store_tile(0, stage);
        if (mine) load_p(0, pc);
        __syncthreads();
        const int nsteps = (nred + 31) / 32;
        for (int st = 0; st < nsteps; ++st) {
            const int buf = st & 1;
            const bool more = st + 1 < nsteps;
            if (more) {
                load_tile(32 * (st + 1), stage);
                if (mine) load_p(32 * (st + 1), pn);
            }
            if (mine) {
                const unsigned char* tb = &tile[buf][r * kTileRowBytes + 16 * q];
#pragma unroll
                for (int x = 0; x < 2; ++x) {
                    bf16x8 hi, lo;
                    split8(pc[x], hi, lo);
#pragma unroll
                    for (int d = 0; d < NB; ++d) {
                        const bf16x8 a = *reinterpret_cast<const bf16x8*>(tb + 16 * d * kTileRowBytes);
                        acc[x][d] = mma(a, hi, acc[x][d]);
                        acc[x][d] = mma(a, lo, acc[x][d]);
                    }
                }
            }
            if (more) {
                store_tile(buf ^ 1, stage);
                if (mine) {
#pragma unroll
                    for (int x = 0; x < 2; ++x)
#pragma unroll
                        for (int v = 0; v < 8; ++v) pc[x][v] = pn[x][v];
                }
            }
            __syncthreads();
        }
    }
    if (o0 >= Lout) return;
    // acc[x][d][v] = grad[row o0 + 16 x + r][16 d + 4 q + v]; zero past the length
#pragma unroll
    for (int x = 0; x < 2; ++x) {
        const int o = o0 + 16 * x + r;
        if (o >= Lout) continue;
        const float g = o < nout ? gv[b] : 0.0f;
        __bf16* dst = out + ((size_t)b * Lout + o) * D + 4 * q;
#pragma unroll
        for (int d = 0; d < NB; ++d)
            *reinterpret_cast<u32x2*>(dst + 16 * d) =
                pack4(g * acc[x][d][0], g * acc[x][d][1], g * acc[x][d][2], g * acc[x][d][3]);
    }
}

// --------------------------------------------------------------------------
// Host side
// --------------------------------------------------------------------------

ohm_status fail(ohm_status st, const std::string& msg) {
    ohm::set_error(msg);
    return st;
}

ohm_status launched(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(OHM_ELAUNCH, std::string(what) + ": " + hipGetErrorString(e));
    return OHM_OK;
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

ohm_status check_dims(const void* Q, const void* K, const int32_t* lengths, int B, int Lq, int Lk,
                      int D) {
    if (B < 0 || Lq < 0 || Lk < 0) return fail(OHM_EINVAL, "B, Lq and Lk must be >= 0");
    if (D != 32 && D != 64 && D != 128 && D != 256)
        return fail(OHM_EINVAL, "embedding dimension D must be 32, 64, 128 or 256");
    if (B > 0 && !lengths) return fail(OHM_EINVAL, "lengths is null");
    if ((int64_t)B * Lq * D > ((int64_t)1 << 40) || (int64_t)B * Lk * D > ((int64_t)1 << 40))
        return fail(OHM_EINVAL, "embedding tensors are too large");
    if (B > 0 && Lq > 0 && (!Q || !aligned16(Q))) return fail(OHM_EINVAL, "Q must be a 16-byte aligned pointer");
    if (B > 0 && Lk > 0 && (!K || !aligned16(K))) return fail(OHM_EINVAL, "K must be a 16-byte aligned pointer");
    return OHM_OK;
}

template <typename F>
ohm_status by_dim(int D, F&& f) {
    switch (D) {
        case 32: return f(std::integral_constant<int, 32>{});
        case 64: return f(std::integral_constant<int, 64>{});
        case 128: return f(std::integral_constant<int, 128>{});
        default: return f(std::integral_constant<int, 256>{});
    }
}

size_t round_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace

extern "C" ohm_status ohm_embed_offsets(const int32_t* lengths, int32_t B, int32_t Lq, int32_t Lk,
                                        int64_t* offsets, void* stream) {
    if (B < 0 || Lq < 0 || Lk < 0) return fail(OHM_EINVAL, "negative batch size or length");
    if (B == 0) return OHM_OK;
    if (!lengths || !offsets) return fail(OHM_EINVAL, "lengths and offsets must be non-null");
    hipLaunchKernelGGL(embed_offsets_kernel, dim3(1), dim3(1024), 0, static_cast<hipStream_t>(stream),
                       lengths, B, Lq, Lk, offsets);
    return launched("embed offsets kernel");
}

extern "C" ohm_status ohm_embed_scores(const void* Q, const void* K, const int32_t* lengths,
                                       int32_t B, int32_t Lq, int32_t Lk, int32_t D,
                                       int64_t* offsets, float* scores, void* stream) {
    ohm_status st = check_dims(Q, K, lengths, B, Lq, Lk, D);
    if (st != OHM_OK) return st;
    if (B == 0) return OHM_OK;
    if (!offsets) return fail(OHM_EINVAL, "offsets is null");
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(embed_offsets_kernel, dim3(1), dim3(1024), 0, s, lengths, B, Lq, Lk, offsets);
    if ((st = launched("embed offsets kernel")) != OHM_OK) return st;
    if (Lq == 0 || Lk == 0) return OHM_OK;
    if (!scores) return fail(OHM_EINVAL, "scores is null");
    const int tiles_i = (Lq + kRows - 1) / kRows;
    const int64_t items = (int64_t)B * tiles_i;
    if (items > (1 << 30)) return fail(OHM_EINVAL, "too many score tiles");
    return by_dim(D, [&](auto dd) {
        constexpr int DD = decltype(dd)::value;
        hipLaunchKernelGGL(embed_scores_kernel<DD>, dim3(round_grid(items)), dim3(kThreads), 0, s,
                           static_cast<const __bf16*>(Q), static_cast<const __bf16*>(K), lengths,
                           offsets, Lq, Lk, tiles_i, (int)items, scores);
        return launched("embed scores kernel");
    });
}

extern "C" size_t ohm_embed_grads_workspace(int32_t B, int32_t Lq, int32_t Lk, int32_t D) {
    if (B <= 0 || D <= 0) return 0;
    const size_t lq = round_up((size_t)(Lq > 0 ? Lq : 0), 32), lk = round_up((size_t)(Lk > 0 ? Lk : 0), 32);
    return round_up((size_t)B * D * lq * 2, 256) + round_up((size_t)B * D * lk * 2, 256);
}

extern "C" ohm_status ohm_embed_grads(const void* Q, const void* K, const int32_t* lengths,
                                      const int64_t* offsets, int32_t B, int32_t Lq, int32_t Lk,
                                      int32_t D, const float* map, const float* grad_value,
                                      void* dQ, void* dK, void* workspace, size_t workspace_bytes,
                                      void* stream) {
    ohm_status st = check_dims(Q, K, lengths, B, Lq, Lk, D);
    if (st != OHM_OK) return st;
    if (B == 0) return OHM_OK;
    if ((Lq > 0 && (!dQ || !aligned16(dQ))) || (Lk > 0 && (!dK || !aligned16(dK))))
        return fail(OHM_EINVAL, "dQ and dK must be 16-byte aligned pointers");
    if (!offsets || !grad_value) return fail(OHM_EINVAL, "offsets and grad_value must be non-null");
    if (Lq > 0 && Lk > 0 && !map) return fail(OHM_EINVAL, "map is null");
    const size_t need = ohm_embed_grads_workspace(B, Lq, Lk, D);
    if (need > 0 && (!workspace || workspace_bytes < need || !aligned16(workspace)))
        return fail(OHM_EWORKSPACE, "workspace missing, misaligned or smaller than ohm_embed_grads_workspace()");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int lq32 = (int)round_up((size_t)Lq, 32), lk32 = (int)round_up((size_t)Lk, 32);
    if ((int64_t)B * ((Lq + kRows - 1) / kRows + (Lk + kRows - 1) / kRows) > (1 << 30))
        return fail(OHM_EINVAL, "too many gradient tiles");
    __bf16* qt = static_cast<__bf16*>(workspace);
    __bf16* kt = reinterpret_cast<__bf16*>(static_cast<char*>(workspace) +
                                          round_up((size_t)B * D * lq32 * 2, 256));
    return by_dim(D, [&](auto dd) {
        constexpr int DD = decltype(dd)::value;
        ohm_status r = OHM_OK;
        if (Lq > 0 && Lk > 0) {
            hipLaunchKernelGGL(embed_transpose_kernel<DD>, dim3((lq32 + 63) / 64, B), dim3(kThreads), 0, s,
                               static_cast<const __bf16*>(Q), lengths, 0, Lq, lq32, qt);
            if ((r = launched("embed transpose kernel (Q)")) != OHM_OK) return r;
            hipLaunchKernelGGL(embed_transpose_kernel<DD>, dim3((lk32 + 63) / 64, B), dim3(kThreads), 0, s,
                               static_cast<const __bf16*>(K), lengths, 1, Lk, lk32, kt);
            if ((r = launched("embed transpose kernel (K)")) != OHM_OK) return r;
        }
        if (Lq > 0) {
            const int tiles = (Lq + kRows - 1) / kRows;
            const int64_t items = (int64_t)B * tiles;
            hipLaunchKernelGGL((embed_grad_kernel<DD, true>), dim3(round_grid(items)), dim3(kThreads), 0, s,
                               kt, map, lengths, offsets, grad_value, Lq, Lk, lk32, tiles, (int)items,
                               static_cast<__bf16*>(dQ));
            if ((r = launched("embed dQ kernel")) != OHM_OK) return r;
        }
        if (Lk > 0) {
            const int tiles = (Lk + kRows - 1) / kRows;
            const int64_t items = (int64_t)B * tiles;
            hipLaunchKernelGGL((embed_grad_kernel<DD, false>), dim3(round_grid(items)), dim3(kThreads), 0, s,
                               qt, map, lengths, offsets, grad_value, Lq, Lk, lq32, tiles, (int)items,
                               static_cast<__bf16*>(dK));
            if ((r = launched("embed dK kernel")) != OHM_OK) return r;
        }
        return r;
    });
}
