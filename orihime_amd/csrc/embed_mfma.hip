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
// A reduction slot run is 16 contiguous bytes (one dwordx4) when the
// reduction index is the fast axis of the array.  For dQ and dK the
// reduction runs over the pair's columns or rows, the slow axis of K and Q,
// so those kernels read transposed copies (Q^T, K^T, zero past each pair's
// length) written by embed_transpose_kernel.  The map P is FP32; it enters
// the BF16 products as two terms, P = hi + lo with hi = bf16(P) and
// lo = bf16(P - hi), which keeps ~16 bits of it.
//
// Work is dealt to the 8 XCDs in contiguous runs (work item w runs on XCD
// w / (grid / 8)), so the tiles of one pair share an L2 and its Q and K rows
// are fetched from HBM about once.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <type_traits>

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

constexpr int kRows = 128;     // output rows per workgroup
constexpr int kThreads = 256;  // four waves (scores: 32 rows each)
constexpr int kGradThreads = 512;  // eight waves (dQ, dK: 16 rows each)
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

// p = hi + lo as two BF16 vectors (v_cvt_pk_bf16_f32: round to nearest even)
DEV void split8(const float (&p)[8], bf16x8& hi, bf16x8& lo) {
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        const __bf16 h = (__bf16)p[v];
        hi[v] = h;
        lo[v] = (__bf16)(p[v] - (float)h);
    }
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
// X [B, L, D] -> XT [B, D, L32] (L32 = L rounded up to 32, so a 32-wide
// reduction run never leaves a row).  Pair b's columns are written up to
// its length rounded up to 32 (zero past the length), the only ones the
// reductions read.
// --------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(kThreads) void embed_transpose_kernel(const __bf16* __restrict__ X,
                                                                   const int* __restrict__ lengths,
                                                                   int side, int L, int L32,
                                                                   __bf16* __restrict__ XT) {
    const int b = blockIdx.y;
    const int len = clampi(lengths[2 * b + side], 0, L);
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    if (j >= ((len + 31) & ~31)) return;
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
// dQ and dK: one workgroup per 128 output rows of a pair, eight waves of 16
// rows x all D.
//   dQ^T[d, i] = sum_j K^T[d, j] P[i, j]   (map rows)
//   dK^T[d, j] = sum_i Q^T[d, i] P[i, j]   (map columns)
// Per reduction step of 32 the D x 32 tile of the transposed copy is shared
// by the eight waves: the workgroup requests it two steps ahead (16 bytes a
// thread) and stores it one step ahead into a ring of three LDS slots (rows
// padded to 80 bytes, so the 16 rows a wave reads at once fall in different
// bank groups); one barrier per step separates a slot's last readers from
// its next writer.  Each wave requests its map values two steps ahead into
// three register sets that trade roles as the loop body is unrolled three
// times (no copies between them).
// Reduction slots: slot v of lane q (= lane / 16) holds reduction index
// 4q + v for v < 4 and 16 + 4q + v - 4 for v >= 4, so the dQ lanes of one
// map row read 64 contiguous bytes per request (two requests per step); the
// LDS tile is stored in the same order.  The order of the products inside
// one MFMA is not the order of the indices; the sums are FP32 either way.
// --------------------------------------------------------------------------
constexpr int kTileRowBytes = 80;  // 64 bytes of a 32-wide BF16 run + padding
constexpr int kTileSlots = 3;

DEV f4u ld16f(const float* base, uint32_t byte_off) {
    return *reinterpret_cast<const f4u*>(reinterpret_cast<const char*>(base) + byte_off);
}
DEV float ld4f(const float* base, uint32_t byte_off) {
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + byte_off);
}
DEV u32x4 ld16x(const __bf16* base, uint32_t byte_off) {
    return *reinterpret_cast<const u32x4*>(reinterpret_cast<const char*>(base) + byte_off);
}

// reduction index of slot v of lane group q (see above)
DEV int red_index(int q, int v) { return v < 4 ? 4 * q + v : 12 + 4 * q + v; }

// LDS position (8-byte units within a tile row) of 8-byte group g of the
// row: group g < 4 (indices 4g .. 4g + 3) at 2g, group g >= 4 at 2(g - 4) + 1
DEV int tile_pos(int g) { return g < 4 ? 2 * g : 2 * (g - 4) + 1; }

template <int N, typename F>
DEV void static_for(F&& f) {
    if constexpr (N > 0) {
        static_for<N - 1>(f);
        f(std::integral_constant<int, N - 1>{});
    }
}

template <int D, bool ROWS>
__global__ __launch_bounds__(kGradThreads) __attribute__((amdgpu_waves_per_eu(D <= 128 ? 4 : 2))) void embed_grad_kernel(
    const __bf16* __restrict__ XT, const float* __restrict__ P, const int* __restrict__ lengths,
    const int64_t* __restrict__ offsets, const float* __restrict__ gv, int Lq, int Lk, int L32,
    int tiles, int items, __bf16* __restrict__ out) {
    constexpr int NB = D / 16;
    constexpr int kPieces = D * 4;  // 16-byte pieces of one D x 32 tile
    constexpr int kPer = (kPieces + kGradThreads - 1) / kGradThreads;
    constexpr int PD = 2;  // map values requested this many steps ahead
    __shared__ __align__(16) unsigned char tile[kTileSlots][D * kTileRowBytes];
    const int w = work_item();
    if (w >= items) return;  // uniform over the workgroup
    const int b = w / tiles, t = w - b * tiles;
    const int M = clampi(lengths[2 * b], 0, Lq), N = clampi(lengths[2 * b + 1], 0, Lk);
    // ROWS: the output rows are the map's rows (dQ), else its columns (dK)
    const int Lout = ROWS ? Lq : Lk, nout = ROWS ? M : N, nred = ROWS ? N : M;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
    const int o0 = t * kRows + 16 * wave;  // this wave's first output row
    const int o = o0 + r;
    f32x4 acc[NB];
#pragma unroll
    for (int d = 0; d < NB; ++d) acc[d] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    // the workgroup runs the reduction when any of its rows is inside the pair
    if (t * kRows < nout && nred > 0) {
        const float* Pb = P + offsets[b];
        const bool mine = o0 < nout;  // this wave has rows inside the pair
        const bool live = o < nout;
        // map values of the step at s0 (wave-uniform bases, 32-bit lane offsets)
        auto load_p = [&](int s0, float (&pv)[8]) {
            if constexpr (ROWS) {
                // row o, columns s0 + red_index(q, v)
                const float* base = Pb + (size_t)o0 * N + s0;
                const uint32_t off = (uint32_t)(r * N + 4 * q) * 4u;
                if (live && s0 + 32 <= nred) {
                    const f4u u0 = ld16f(base, off), u1 = ld16f(base, off + 64);
#pragma unroll
                    for (int v = 0; v < 4; ++v) { pv[v] = u0[v]; pv[4 + v] = u1[v]; }
                } else {
#pragma unroll
                    for (int v = 0; v < 8; ++v) {
                        const int j = red_index(q, v);
                        pv[v] = (live && s0 + j < nred) ? ld4f(base, off + 4u * (uint32_t)(j - 4 * q)) : 0.0f;
                    }
                }
            } else {
                // column o, rows s0 + red_index(q, v)
                const float* base = Pb + (size_t)s0 * N + o0;
                const uint32_t off = (uint32_t)(4 * q * N + r) * 4u;
#pragma unroll
                for (int v = 0; v < 8; ++v) {
                    const int j = red_index(q, v);
                    const float* bv = base + (size_t)(j - 4 * q) * N;
                    pv[v] = (live && s0 + j < nred) ? ld4f(bv, off) : 0.0f;
                }
            }
        };
        // the workgroup's share of one D x 32 tile: piece c is row c / 4,
        // bytes 16 (c % 4) of the row's 64, i.e. 8-byte groups 2 (c % 4) and
        // 2 (c % 4) + 1
        const __bf16* xt = XT + (size_t)b * D * L32;
        auto load_tile = [&](int s0, u32x4 (&stg)[kPer]) {
            const __bf16* base = xt + s0;
#pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const int c = threadIdx.x + kGradThreads * k;
                if (c < kPieces) stg[k] = ld16x(base, (uint32_t)((c >> 2) * L32 + 8 * (c & 3)) * 2u);
            }
        };
        auto store_tile = [&](int slot, const u32x4 (&stg)[kPer]) {
#pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const int c = threadIdx.x + kGradThreads * k;
                if (c < kPieces) {
                    unsigned char* row = &tile[slot][(c >> 2) * kTileRowBytes];
                    const int g = 2 * (c & 3);
                    *reinterpret_cast<u32x2*>(row + 8 * tile_pos(g)) = u32x2{stg[k][0], stg[k][1]};
                    *reinterpret_cast<u32x2*>(row + 8 * tile_pos(g + 1)) = u32x2{stg[k][2], stg[k][3]};
                }
            }
        };
        const int nsteps = (nred + 31) / 32;
        u32x4 stage[kPer];
        float ps[PD + 1][8];
        load_tile(0, stage);
        store_tile(0, stage);
        if (nsteps > 1) {
            load_tile(32, stage);
            store_tile(1, stage);
        }
        if (nsteps > 2) load_tile(64, stage);
        if (mine) {
            static_for<PD>([&](auto J) {
                if (J < nsteps) load_p(32 * J, ps[J]);
            });
        }
        __syncthreads();
        // step st: products from slot st % 3 and `cur`; the tile of step
        // st + 2 goes into its slot (whose last readers passed the barrier
        // of step st - 1), the tile of step st + 3 is requested, and the map
        // values of step st + 2 go into `nxt`
        auto body = [&](int st, const float (&cur)[8], float (&nxt)[8]) {
            if (st + 2 < nsteps) store_tile((st + 2) % kTileSlots, stage);
            if (st + 3 < nsteps) load_tile(32 * (st + 3), stage);
            if (mine) {
                if (st + PD < nsteps) load_p(32 * (st + PD), nxt);
                bf16x8 hi, lo;
                split8(cur, hi, lo);
                const unsigned char* tb = &tile[st % kTileSlots][r * kTileRowBytes + 16 * q];
#pragma unroll
                for (int d = 0; d < NB; ++d) {
                    const bf16x8 a = *reinterpret_cast<const bf16x8*>(tb + 16 * d * kTileRowBytes);
                    acc[d] = mma(a, hi, acc[d]);
                    acc[d] = mma(a, lo, acc[d]);
                }
            }
            __syncthreads();
        };
        for (int st = 0; st < nsteps; st += PD + 1) {
            static_for<PD + 1>([&](auto J) {
                constexpr int j = decltype(J)::value;
                if (st + j < nsteps) body(st + j, ps[j], ps[(j + PD) % (PD + 1)]);
            });
        }
    }
    if (o >= Lout) return;
    // acc[d][v] = grad[row o][16 d + 4 q + v]; zero past the length
    const float g = o < nout ? gv[b] : 0.0f;
    __bf16* dst = out + ((size_t)b * Lout + o) * D + 4 * q;
#pragma unroll
    for (int d = 0; d < NB; ++d)
        *reinterpret_cast<u32x2*>(dst + 16 * d) =
            pack4(g * acc[d][0], g * acc[d][1], g * acc[d][2], g * acc[d][3]);
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
            hipLaunchKernelGGL((embed_grad_kernel<DD, true>), dim3(round_grid(items)), dim3(kGradThreads), 0, s,
                               kt, map, lengths, offsets, grad_value, Lq, Lk, lk32, tiles, (int)items,
                               static_cast<__bf16*>(dQ));
            if ((r = launched("embed dQ kernel")) != OHM_OK) return r;
        }
        if (Lk > 0) {
            const int tiles = (Lk + kRows - 1) / kRows;
            const int64_t items = (int64_t)B * tiles;
            hipLaunchKernelGGL((embed_grad_kernel<DD, false>), dim3(round_grid(items)), dim3(kGradThreads), 0, s,
                               qt, map, lengths, offsets, grad_value, Lq, Lk, lq32, tiles, (int)items,
                               static_cast<__bf16*>(dK));
            if ((r = launched("embed dK kernel")) != OHM_OK) return r;
        }
        return r;
    });
}
