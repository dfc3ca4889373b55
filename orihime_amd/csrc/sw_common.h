// Shared definitions for the soft Smith-Waterman library (host and device).
//
// Units.  Every DP value inside the kernels is kept in log2 units scaled by
// the temperature: x~ = x * log2(e) / T.  A temperature-scaled log-sum-exp
// LSE_T(x_k) = T ln sum exp(x_k / T) then becomes log2(sum 2^{x~_k}), so the
// inner loops use the hardware exp2/log2 and never divide by T.  Values are
// converted back to natural units (multiply by T ln 2) only for outputs.
//
// Recurrence (affine).  Following src/sw_affine/kernels_cpu.cpp:94-195 the
// cell values are
//   M(i,j) = LSE_T(M'+s, I'+s, D'+s, s)        (' = cell (i-1, j-1), i,j > 1)
//   I(i,j) = LSE_T(M(i-1,j)+go, I(i-1,j)+ge)   (i > 1)
//   D(i,j) = LSE_T(M(i,j-1)+go, D(i,j-1)+ge)   (j > 1)
//   S      = LSE_T(0, {M, I, D at every cell})
// which this library evaluates through the per-cell total A = LSE_T(M, I, D):
//   M(i,j) = s + LSE_T(A(i-1,j-1), 0),   S = LSE_T(0, {A at every cell}).
// The two forms are the same function of the inputs (nested log-sum-exp at
// one temperature is associative), so values and every derivative agree.
//
// Linear: A(i,j) = LSE_T(A'+s, A(i-1,j)+g, A(i,j-1)+g, s)
//                = LSE_T(X, A(i-1,j)+g, A(i,j-1)+g),  X = s + LSE_T(A', 0)
// (src/sw/kernels_cpu.cpp:70-148).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <string>

namespace ohm {

// Records the message returned by ohm_last_error() on this thread.
void set_error(const std::string& msg);

constexpr float kNinf = -1e30f;             // "minus infinity" of the reference, numerics.h:29
constexpr float kLog2e = 1.4426950408889634f;
constexpr float kLn2 = 0.6931471805599453f;
constexpr int kWave = 64;                   // gfx950 wavefront
constexpr int kStrip = 64;                  // rows per strip = lanes per wave
constexpr int kChunk = 16;                  // columns per unrolled chunk
constexpr int kBndStride = 8;               // floats per boundary entry (one column):
                                            // [0..3] forward sweep, [4..7] reverse sweep

// Per-pair work descriptor built on the device by the plan kernel.
struct PairDesc {
    int64_t s_off;  // element offset of the pair's scores (and map) block
    int64_t t_off;  // element offset of the pair's DP table
    int64_t b_off;  // element offset of the pair's strip-boundary block
    int32_t ld;     // row stride of scores / map
    int32_t L1, L2; // active lengths
    int32_t ns;     // strips = ceil(L1 / 64) (0 when the pair has no cells)
    int32_t lost;   // 1: the workspace had no room for the pair (it runs as an
                    // empty pair and its outputs are overwritten with NaN)
};

// Floats per table step (one column of a strip) for each pass: the table of
// a strip is [L2 + 63 steps][planes][64 lanes].
// What the forward sweep keeps for the reverse sweep, per strip and per
// stored step: `planes` rows of 64 lanes every `stride` anti-diagonals.
//   affine: the sweep state every kCkpt steps (checkpoints; the reverse sweep
//           recomputes the steps in between),
//   linear and the tangent passes: every step (full table).
#ifndef OHM_CKPT
#define OHM_CKPT 8
#endif
constexpr int kCkpt = OHM_CKPT;         // steps between affine checkpoints (divides kChunk)
constexpr int kPlanesAffineCkpt = 6;    // six state rows (log-domain sweep: M, I, D, A of the
                                        // previous step and A of the diagonal; scaled sweep:
                                        // g, h, a, c of the previous step, a and c of the diagonal)
constexpr int kPlanesAffineTan = 7;     // M, I, D and their tangents (log-domain tangent sweeps),
                                        // plus the evaluation exponent (exponent-tracked HVP)
constexpr int kPlanesLinearFwd = 2;     // A, X
constexpr int kPlanesLinearTan = 4;     // A, X and their tangents

struct TableSpec {
    int planes, stride;
};

// Upper bound of sum_b ns_b * ceil((L2_b + 63) / stride) * 64 * planes over a
// batch with N cells in total (ns * 64 <= L1 + 63).
inline int64_t table_bound(int64_t N, int64_t B, int64_t L1max, int64_t L2max, TableSpec t) {
    const int64_t full = N + 63 * B * (L1max + L2max) + 3969 * B;  // sum (L1+63)(L2+63)
    return (int64_t)t.planes * ((full + t.stride - 1) / t.stride + B * (L1max + 63));
}
inline int64_t bnd_bound(int64_t N, int64_t B, int64_t L2max) {
    return (int64_t)kBndStride * (N / 64 + 1 + B * L2max);
}

// Workspace layout, byte offsets, 256-byte aligned.  Each pair's DP table
// and its strip-boundary block are contiguous inside one data region whose
// offsets the plan kernel computes exactly; `data` bytes bound the region.
struct WsLayout {
    size_t claims, desc, order, stil, tgt, pst, grid, data, total;
};

// Per-pair state words of the scaled sweeps (4 int32 per pair).
constexpr int kPstFlag = 0;   // 1: the scaled sweep left its range, the log-domain sweep redoes the pair
constexpr int kPstZexp = 1;   // partition Z = zm * 2^zexp, zm in [1, 2)
constexpr int kPstZinv = 2;   // bits of 1 / zm
// Slots of the claims block (64 ints, zeroed by the plan kernel): pair claim
// counters of the launches, and the count of flagged pairs.
constexpr int kClaimFlagged = 8;
constexpr int kClaimGridFwd = 10;   // strip tickets of the grid-scheduled sweeps
constexpr int kClaimGridRev = 11;
constexpr int kClaimGridStall = 12;  // set when a grid hand-off wait exceeded its bound
constexpr int kClaimFused = 13;      // pair claims of the fused forward + reverse launch
constexpr int kClaimWsShort = 14;    // set by the plan kernel when some pair did not fit the workspace

// Grid scheduling (strips of a pair on waves of any workgroup, hand-off
// through global memory) for batches of at most kGridMaxPairs pairs.  Its
// region: per-pair strip completion counts of the forward and the reverse
// launch [2][B] (zeroed by the plan kernel), then per-strip partial results
// [B][max_strips][4] floats.
constexpr int kGridMaxPairs = 64;
inline int64_t grid_zero_ints(int64_t B, int64_t) { return 2 * B; }
inline size_t grid_bytes(int64_t B, int64_t max_strips) {
    if (B > kGridMaxPairs) return 0;
    return (size_t)(grid_zero_ints(B, max_strips) + B * max_strips * 4) * 4;
}

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

inline size_t ws_fixed(int64_t B, int64_t max_strips, size_t* claims, size_t* desc, size_t* order,
                       size_t* stil, size_t* tgt, size_t* pst, size_t* grid) {
    size_t o = 0;
    *claims = o; o += 256;
    *desc = o; o = align256(o + (size_t)B * sizeof(PairDesc));
    *order = o; o = align256(o + (size_t)B * sizeof(int32_t));
    *stil = o; o = align256(o + (size_t)B * sizeof(float));
    *tgt = o; o = align256(o + (size_t)B * 4 * sizeof(float));
    *pst = o; o = align256(o + (size_t)B * 4 * sizeof(int32_t));
    *grid = o; o = align256(o + grid_bytes(B, max_strips));
    return o;
}

// op: 0 value only, 1 forward + backward, 2 tangent passes.
inline WsLayout ws_layout(int op, int64_t B, int64_t L1max, int64_t L2max, int64_t N,
                          TableSpec fwd, TableSpec tan) {
    if (N < 0) N = B * L1max * L2max;
    WsLayout w;
    const int64_t max_strips = L1max > 0 ? (L1max + kStrip - 1) / kStrip : 1;
    size_t o = ws_fixed(B, max_strips, &w.claims, &w.desc, &w.order, &w.stil, &w.tgt, &w.pst,
                        &w.grid);
    w.data = o;
    const TableSpec none = {0, 1};
    const TableSpec t = op == 0 ? none : (op == 1 ? fwd : tan);
    // + 64 floats per pair: each pair's region is rounded up to 256 bytes
    int64_t elems = table_bound(N, B, L1max, L2max, t) + bnd_bound(N, B, L2max) + 64 * B;
    w.total = align256(o + (size_t)elems * sizeof(float));
    return w;
}

}  // namespace ohm
