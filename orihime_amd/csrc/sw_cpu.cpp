// Host (CPU) implementation of the soft Smith-Waterman entry points, for CPU
// tensors — the counterpart of the reference's CPU dispatch
// (src/sw_affine/torch_cpu.cpp:593-707, src/sw/torch_cpu.cpp:522-591).
//
// Same mathematics as the GPU kernels (sw_kernels.hip): log2-scaled values,
// the per-cell total A = LSE(M, I, D), a pull-form adjoint and
// forward-over-reverse tangents.  Each pair is swept row-major (forward) and
// reverse row-major (adjoint), which visits every dependency before use;
// pairs run in parallel on std::threads.  Per-pair sums are accumulated in
// double, so results do not depend on the thread count.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "ohm_sw.h"
#include "sw_common.h"

using namespace ohm;

namespace {

ohm_status cpu_fail(ohm_status st, const char* msg) {
    ohm::set_error(msg);
    return st;
}

inline float ex2(float x) { return std::exp2(x); }
inline float lg2(float x) { return std::log2(x); }
inline float clampw(float l) { return std::max(std::min(l, 0.0f), -256.0f); }

// y = log2(2^a + 2^b), weights, entropy term h = y - (wa a + wb b).
inline void lse2(float a, float b, float& y, float& wa, float& wb, float& h) {
    float hi = std::max(a, b), lo = std::min(a, b);
    float e = ex2(lo - hi), r = 1.0f / (1.0f + e), l1 = lg2(1.0f + e), wl = e * r;
    y = hi + l1;
    bool a_hi = a >= b;
    wa = a_hi ? r : wl;
    wb = a_hi ? wl : r;
    h = l1 + wl * std::min(hi - lo, 512.0f);
}
// y = log2(2^a + 2^b + 2^c) with weights normalised to sum to 1 (they are
// divided by their own sum, not taken as 2^(x - y) with a rounded y, so no
// flow is lost through the node) and entropy term h = y - sum w x.
inline void lse3(float a, float b, float c, float& y, float& wa, float& wb, float& wc, float& h) {
    float hi = std::max(std::max(a, b), c);
    float ea = ex2(a - hi), eb = ex2(b - hi), ec = ex2(c - hi);
    float sum = ea + eb + ec, r = 1.0f / sum, l = lg2(sum);
    y = hi + l;
    wa = ea * r; wb = eb * r; wc = ec * r;
    h = l + wa * std::min(hi - a, 512.0f) + wb * std::min(hi - b, 512.0f) + wc * std::min(hi - c, 512.0f);
}
inline void softplus(float x, float& y, float& w, float& h) {
    float e = ex2(-std::fabs(x)), r = 1.0f / (1.0f + e), l1 = lg2(1.0f + e), ws = e * r;
    y = std::max(x, 0.0f) + l1;
    w = x >= 0.0f ? r : ws;
    h = l1 + ws * std::min(std::fabs(x), 512.0f);
}
inline void pair_w(float la, float lb, float& wa, float& wb) {
    float lo = std::min(std::min(la, lb), 0.0f), wl = ex2(lo), wh = 1.0f - wl;
    bool a_lo = la <= lb;
    wa = a_lo ? wl : wh;
    wb = a_lo ? wh : wl;
}
inline float sigmoid(float x) { return 1.0f / (1.0f + ex2(-x)); }
// 1 - sigmoid(x), without cancellation
inline float sigmoid_c(float x) { return 1.0f / (1.0f + ex2(x)); }

struct Problem {
    const float* scores;
    const float* dir;
    const int32_t* lengths;
    const int64_t* offsets;
    int B, L1max, L2max;
    float go, ge, T;          // natural units; linear uses go as the gap
    float kap, go2, ge2;      // log2 units
    float dgo, dge, dT;       // tangent directions
    float* value;
    float* map;
    float* map_t;
    float* g0;
    float* g1;
    float* g2;
    float* contract;
    bool gap_sum;             // NW / SV: g0 = opens + extends (linear-gap models)
};

struct PairRef {
    int64_t off;
    int ld, L1, L2;
};

PairRef pair_ref(const Problem& P, int b) {
    PairRef r;
    int L1 = P.lengths ? P.lengths[2 * b] : P.L1max;
    int L2 = P.lengths ? P.lengths[2 * b + 1] : P.L2max;
    r.L1 = L1;
    r.L2 = L2;
    r.off = P.offsets ? P.offsets[b] : (int64_t)b * P.L1max * P.L2max;
    r.ld = P.offsets ? L2 : P.L2max;
    return r;
}

// Partition accumulation in double: returns S~ (log2 units) and, with
// tangents, S' (natural units).
struct Part {
    double sum(double& dS, const std::vector<float>& A, const std::vector<float>* dA, double dTl) {
        double m = 0.0;  // empty alignment
        for (float v : A) if (v > kNinf) m = std::max(m, (double)v);
        double z = std::exp2(-m), za = 0.0, y = std::exp2(-m) * (-m);
        for (size_t q = 0; q < A.size(); ++q) {
            if (!(A[q] > kNinf)) continue;
            double w = std::exp2((double)A[q] - m);
            z += w;
            if (dA) za += w * (*dA)[q];
            y += w * ((double)A[q] - m);
        }
        double lz = std::log2(z);
        dS = za / z + dTl * (lz - y / z);  // dTl = d_temp * ln 2
        return m + lz;
    }
};

void zero_pad(const Problem& P, const PairRef& r, float* out) {
    if (!out || P.offsets) return;
    for (int i = 0; i < P.L1max; ++i)
        for (int j = 0; j < P.L2max; ++j)
            if (i >= r.L1 || j >= r.L2) out[r.off + (int64_t)i * P.L2max + j] = 0.0f;
}

// --------------------------------------------------------------------------
// Affine: tangent-capable forward and adjoint for one pair.  With TAN false
// the tangent arithmetic is skipped.
// --------------------------------------------------------------------------

template <bool TAN>
void affine_pair(const Problem& P, int b, bool want_bwd) {
    const PairRef r = pair_ref(P, b);
    const int L1 = r.L1, L2 = r.L2, C = L2 + 1;
    const size_t n = (size_t)(L1 + 1) * C;
    std::vector<float> M(n, kNinf), I(n, kNinf), D(n, kNinf), A(n, kNinf);
    std::vector<float> dM, dI, dD, dA;
    if (TAN) { dM.assign(n, 0.f); dI.assign(n, 0.f); dD.assign(n, 0.f); dA.assign(n, 0.f); }
    const float* s = P.scores + r.off;
    const float* v = P.dir ? P.dir + r.off : nullptr;
    const float dTl = P.dT * kLn2;
    auto at = [C](int i, int j) { return (size_t)i * C + j; };
    for (int i = 1; i <= L1; ++i)
        for (int j = 1; j <= L2; ++j) {
            const size_t c = at(i, j), dg = at(i - 1, j - 1), up = at(i - 1, j), lf = at(i, j - 1);
            float sp, wA, hA;
            softplus(A[dg], sp, wA, hA);
            M[c] = s[(int64_t)(i - 1) * r.ld + (j - 1)] * P.kap + sp;
            float y, w1, w2, h;
            if (i > 1) {
                lse2(M[up] + P.go2, I[up] + P.ge2, y, w1, w2, h);
                I[c] = y;
                if (TAN) dI[c] = w1 * (dM[up] + P.dgo) + w2 * (dI[up] + P.dge) + dTl * h;
            }
            if (j > 1) {
                lse2(M[lf] + P.go2, D[lf] + P.ge2, y, w1, w2, h);
                D[c] = y;
                if (TAN) dD[c] = w1 * (dM[lf] + P.dgo) + w2 * (dD[lf] + P.dge) + dTl * h;
            }
            float w3;
            lse3(M[c], I[c], D[c], y, w1, w2, w3, h);
            A[c] = y;
            if (TAN) {
                dM[c] = (v ? v[(int64_t)(i - 1) * r.ld + (j - 1)] : 0.0f) + wA * dA[dg] + dTl * hA;
                dA[c] = w1 * dM[c] + w2 * dI[c] + w3 * dD[c] + dTl * h;
            }
        }
    // (0,0) carries the empty alignment only through the partition
    double dS = 0.0;
    std::vector<float> Acells;
    Acells.reserve((size_t)L1 * L2);
    std::vector<float> dAcells;
    for (int i = 1; i <= L1; ++i)
        for (int j = 1; j <= L2; ++j) {
            Acells.push_back(A[at(i, j)]);
            if (TAN) dAcells.push_back(dA[at(i, j)]);
        }
    Part part;
    const double St = part.sum(dS, Acells, TAN ? &dAcells : nullptr, dTl);
    if (P.value) P.value[b] = (float)(St / P.kap);
    if (!want_bwd) return;

    // Adjoint (and its tangent), reverse row-major.
    const float Stf = (float)St, dSf = (float)dS, iT = 1.0f / P.T;
    std::vector<float> bM(n, 0.f), bI(n, 0.f), bD(n, 0.f), dbM, dbI, dbD;
    if (TAN) { dbM.assign(n, 0.f); dbI.assign(n, 0.f); dbD.assign(n, 0.f); }
    double acc0 = 0, acc1 = 0, acc2 = 0, tac0 = 0, tac1 = 0, tac2 = 0;
    for (int i = L1; i >= 1; --i)
        for (int j = L2; j >= 1; --j) {
            const size_t c = at(i, j);
            const bool has_i = i < L1, has_d = j < L2, has_m = has_i && has_d;
            float y, wM, wI, wD, hA;
            lse3(M[c], I[c], D[c], y, wM, wI, wD, hA);
            const float Ac = A[c];
            const float lt = clampw(Ac - Stf), pi = ex2(lt), sg = sigmoid(Ac);
            const float bMs = has_m ? bM[at(i + 1, j + 1)] : 0.f;
            float bA = pi + bMs * sg;
            float tIM = 0, tII = 0, tDM = 0, tDD = 0, wIM = 0, wII = 0, wDM = 0, wDD = 0;
            float lIM = 0, lII = 0, lDM = 0, lDD = 0;
            if (has_i) {
                const size_t q = at(i + 1, j);
                lIM = clampw(M[c] + P.go2 - I[q]);
                lII = clampw(I[c] + P.ge2 - I[q]);
                pair_w(lIM, lII, wIM, wII);
                tIM = bI[q] * wIM;
                tII = bI[q] * wII;
            }
            if (has_d) {
                const size_t q = at(i, j + 1);
                lDM = clampw(M[c] + P.go2 - D[q]);
                lDD = clampw(D[c] + P.ge2 - D[q]);
                pair_w(lDM, lDD, wDM, wDD);
                tDM = bD[q] * wDM;
                tDD = bD[q] * wDD;
            }
            bM[c] = bA * wM + tIM + tDM;
            bI[c] = bA * wI + tII;
            bD[c] = bA * wD + tDD;
            const float sc = s[(int64_t)(i - 1) * r.ld + (j - 1)];
            acc0 += (double)tIM + tDM;
            acc1 += (double)tII + tDD;
            acc2 += (double)bM[c] * sc;
            if (P.map) P.map[r.off + (int64_t)(i - 1) * r.ld + (j - 1)] = bM[c];
            if (TAN) {
                const float dAc = dA[c];
                const float dpi = pi * ((dAc - dSf) - dTl * lt) * iT;
                const float dsg = sg * sigmoid_c(Ac) * (dAc - dTl * std::max(Ac, -256.0f)) * iT;
                const float dbMs = has_m ? dbM[at(i + 1, j + 1)] : 0.f;
                const float dbA = dpi + dbMs * sg + bMs * dsg;
                const float dwM = wM * ((dM[c] - dAc) - dTl * clampw(M[c] - Ac)) * iT;
                const float dwI = wI * ((dI[c] - dAc) - dTl * clampw(I[c] - Ac)) * iT;
                const float dwD = wD * ((dD[c] - dAc) - dTl * clampw(D[c] - Ac)) * iT;
                float dtIM = 0, dtII = 0, dtDM = 0, dtDD = 0;
                if (has_i) {
                    const size_t q = at(i + 1, j);
                    const float dwIM = wIM * ((dM[c] + P.dgo - dI[q]) - dTl * lIM) * iT;
                    const float dwII = wII * ((dI[c] + P.dge - dI[q]) - dTl * lII) * iT;
                    dtIM = dbI[q] * wIM + bI[q] * dwIM;
                    dtII = dbI[q] * wII + bI[q] * dwII;
                }
                if (has_d) {
                    const size_t q = at(i, j + 1);
                    const float dwDM = wDM * ((dM[c] + P.dgo - dD[q]) - dTl * lDM) * iT;
                    const float dwDD = wDD * ((dD[c] + P.dge - dD[q]) - dTl * lDD) * iT;
                    dtDM = dbD[q] * wDM + bD[q] * dwDM;
                    dtDD = dbD[q] * wDD + bD[q] * dwDD;
                }
                dbM[c] = dbA * wM + bA * dwM + dtIM + dtDM;
                dbI[c] = dbA * wI + bA * dwI + dtII;
                dbD[c] = dbA * wD + bA * dwD + dtDD;
                tac0 += (double)dtIM + dtDM;
                tac1 += (double)dtII + dtDD;
                tac2 += (double)dbM[c] * sc;
                if (P.map_t) P.map_t[r.off + (int64_t)(i - 1) * r.ld + (j - 1)] = dbM[c];
            }
        }
    const double S = St / P.kap;
    if (P.g0) P.g0[b] = (float)acc0;
    if (P.g1) P.g1[b] = (float)acc1;
    if (P.g2) P.g2[b] = (float)((S - (acc2 + (double)P.go * acc0 + (double)P.ge * acc1)) / P.T);
    if (TAN && P.contract) {
        P.contract[3 * b + 0] = (float)tac0;
        P.contract[3 * b + 1] = (float)tac1;
        P.contract[3 * b + 2] = (float)(-(tac2 + (double)P.go * tac0 + (double)P.ge * tac1) / P.T);
    }
    zero_pad(P, r, P.map);
    zero_pad(P, r, P.map_t);
}

// --------------------------------------------------------------------------
// Linear gap
// --------------------------------------------------------------------------

template <bool TAN>
void linear_pair(const Problem& P, int b, bool want_bwd) {
    const PairRef r = pair_ref(P, b);
    const int L1 = r.L1, L2 = r.L2, C = L2 + 1;
    const size_t n = (size_t)(L1 + 1) * C;
    std::vector<float> A(n, kNinf), X(n, kNinf), dA, dX;
    if (TAN) { dA.assign(n, 0.f); dX.assign(n, 0.f); }
    const float* s = P.scores + r.off;
    const float* v = P.dir ? P.dir + r.off : nullptr;
    const float dTl = P.dT * kLn2;
    auto at = [C](int i, int j) { return (size_t)i * C + j; };
    for (int i = 1; i <= L1; ++i)
        for (int j = 1; j <= L2; ++j) {
            const size_t c = at(i, j), dg = at(i - 1, j - 1), up = at(i - 1, j), lf = at(i, j - 1);
            float sp, wA, hA;
            softplus(A[dg], sp, wA, hA);
            X[c] = s[(int64_t)(i - 1) * r.ld + (j - 1)] * P.kap + sp;
            const float cu = i > 1 ? A[up] + P.go2 : kNinf;
            const float cl = j > 1 ? A[lf] + P.go2 : kNinf;
            float y, wx, wu, wl, h;
            lse3(X[c], cu, cl, y, wx, wu, wl, h);
            A[c] = y;
            if (TAN) {
                dX[c] = (v ? v[(int64_t)(i - 1) * r.ld + (j - 1)] : 0.0f) + wA * dA[dg] + dTl * hA;
                dA[c] = wx * dX[c] + wu * (dA[up] + P.dgo) + wl * (dA[lf] + P.dgo) + dTl * h;
            }
        }
    double dS = 0.0;
    std::vector<float> Acells, dAcells;
    for (int i = 1; i <= L1; ++i)
        for (int j = 1; j <= L2; ++j) {
            Acells.push_back(A[at(i, j)]);
            if (TAN) dAcells.push_back(dA[at(i, j)]);
        }
    Part part;
    const double St = part.sum(dS, Acells, TAN ? &dAcells : nullptr, dTl);
    if (P.value) P.value[b] = (float)(St / P.kap);
    if (!want_bwd) return;

    // Adjoint in edge-flow form: node c splits its adjoint over its three
    // in-edges with weights normalised at c; the flows fu (edge from
    // (i-1, j)) and fl (edge from (i, j-1)) are pulled by the predecessors.
    const float Stf = (float)St, dSf = (float)dS, iT = 1.0f / P.T;
    std::vector<float> gm(n, 0.f), fu(n, 0.f), fl(n, 0.f), dgm, dfu, dfl;
    if (TAN) { dgm.assign(n, 0.f); dfu.assign(n, 0.f); dfl.assign(n, 0.f); }
    double acc0 = 0, acc1 = 0, tac0 = 0, tac1 = 0;
    for (int i = L1; i >= 1; --i)
        for (int j = L2; j >= 1; --j) {
            const size_t c = at(i, j), up = at(i - 1, j), lf = at(i, j - 1);
            const bool has_i = i < L1, has_d = j < L2, has_m = has_i && has_d;
            const float Ac = A[c];
            const float cu = i > 1 ? A[up] + P.go2 : kNinf;
            const float cl = j > 1 ? A[lf] + P.go2 : kNinf;
            float y, wx, wu, wl, h;
            lse3(X[c], cu, cl, y, wx, wu, wl, h);
            const float lt = clampw(Ac - Stf), pi = ex2(lt), sg = sigmoid(Ac);
            const float gD = has_m ? gm[at(i + 1, j + 1)] : 0.f;
            const float inU = has_i ? fu[at(i + 1, j)] : 0.f;
            const float inL = has_d ? fl[at(i, j + 1)] : 0.f;
            const float bt = pi + gD * sg + inU + inL;
            gm[c] = bt * wx;
            fu[c] = bt * wu;
            fl[c] = bt * wl;
            const float sc = s[(int64_t)(i - 1) * r.ld + (j - 1)];
            acc0 += (double)fu[c] + fl[c];
            acc1 += (double)gm[c] * sc;
            if (P.map) P.map[r.off + (int64_t)(i - 1) * r.ld + (j - 1)] = gm[c];
            if (TAN) {
                const float dAc = dA[c];
                const float dpi = pi * ((dAc - dSf) - dTl * lt) * iT;
                const float dsg = sg * sigmoid_c(Ac) * (dAc - dTl * std::max(Ac, -256.0f)) * iT;
                const float dgD = has_m ? dgm[at(i + 1, j + 1)] : 0.f;
                const float dinU = has_i ? dfu[at(i + 1, j)] : 0.f;
                const float dinL = has_d ? dfl[at(i, j + 1)] : 0.f;
                const float dbt = dpi + dgD * sg + gD * dsg + dinU + dinL;
                const float dwx = wx * ((dX[c] - dAc) - dTl * clampw(X[c] - Ac)) * iT;
                const float dwu = i > 1 ? wu * ((dA[up] + P.dgo - dAc) - dTl * clampw(cu - Ac)) * iT : 0.f;
                const float dwl = j > 1 ? wl * ((dA[lf] + P.dgo - dAc) - dTl * clampw(cl - Ac)) * iT : 0.f;
                dgm[c] = dbt * wx + bt * dwx;
                dfu[c] = dbt * wu + bt * dwu;
                dfl[c] = dbt * wl + bt * dwl;
                tac0 += (double)dfu[c] + dfl[c];
                tac1 += (double)dgm[c] * sc;
                if (P.map_t) P.map_t[r.off + (int64_t)(i - 1) * r.ld + (j - 1)] = dgm[c];
            }
        }
    const double S = St / P.kap;
    if (P.g0) P.g0[b] = (float)acc0;
    if (P.g1) P.g1[b] = (float)((S - (acc1 + (double)P.go * acc0)) / P.T);
    if (TAN && P.contract) {
        P.contract[2 * b + 0] = (float)tac0;
        P.contract[2 * b + 1] = (float)(-(tac1 + (double)P.go * tac0) / P.T);
    }
    zero_pad(P, r, P.map);
    zero_pad(P, r, P.map_t);
}

// --------------------------------------------------------------------------
// Needleman-Wunsch, affine gap (global alignment), one pair:
// src/nw_affine/kernels_cpu.cpp:100-335.  Full M, I, D tables in log2 units
// including the boundary row and column, then the pull-form adjoint in
// reverse row-major order: the mass of each state is split over its three
// predecessor states (all in one cell) by weights normalised at that cell.
// --------------------------------------------------------------------------

void nw_pair(const Problem& P, int b, bool want_bwd) {
    const PairRef r = pair_ref(P, b);
    const int L1 = r.L1, L2 = r.L2;
    if (L1 == 0 || L2 == 0) {
        // one leading gap over the non-empty side, or the empty alignment
        const int n = L1 + L2;
        P.value[b] = n == 0 ? 0.0f : P.go + (float)(n - 1) * P.ge;
        if (P.g0) P.g0[b] = n > 0 ? (P.gap_sum ? (float)n : 1.0f) : 0.0f;
        if (P.g1) P.g1[b] = n > 0 ? (float)(n - 1) : 0.0f;
        if (P.g2) P.g2[b] = 0.0f;
        if (want_bwd) zero_pad(P, r, P.map);
        return;
    }
    const float* s = P.scores + r.off;
    const int C = L2 + 1;
    const size_t n = (size_t)(L1 + 1) * C;
    auto at = [C](int i, int j) { return (size_t)i * C + j; };
    auto l2 = [](float a, float c) {
        float hi = std::max(a, c), lo = std::min(a, c);
        return hi + lg2(1.0f + ex2(lo - hi));
    };
    auto l3 = [&](float a, float c, float d) { return l2(l2(a, c), d); };
    std::vector<float> M(n, kNinf), I(n, kNinf), D(n, kNinf);
    M[0] = 0.0f;
    for (int i = 1; i <= L1; ++i) I[at(i, 0)] = P.go2 + (float)(i - 1) * P.ge2;
    for (int j = 1; j <= L2; ++j) D[at(0, j)] = P.go2 + (float)(j - 1) * P.ge2;
    for (int i = 1; i <= L1; ++i)
        for (int j = 1; j <= L2; ++j) {
            const size_t c = at(i, j), dg = at(i - 1, j - 1), up = at(i - 1, j), lf = at(i, j - 1);
            M[c] = s[(int64_t)(i - 1) * r.ld + j - 1] * P.kap + l3(M[dg], I[dg], D[dg]);
            I[c] = l3(M[up] + P.go2, I[up] + P.ge2, D[up] + P.go2);
            D[c] = l3(M[lf] + P.go2, I[lf] + P.go2, D[lf] + P.ge2);
        }
    const size_t f = at(L1, L2);
    const double S = (double)l3(M[f], I[f], D[f]) / P.kap;
    P.value[b] = (float)S;
    if (!want_bwd) return;
    auto softmax = [](float a, float c, float d, float* w) {
        float hi = std::max(std::max(a, c), d);
        float ea = ex2(a - hi), ec = ex2(c - hi), ed = ex2(d - hi), q = 1.0f / (ea + ec + ed);
        w[0] = ea * q; w[1] = ec * q; w[2] = ed * q;
    };
    std::vector<float> PM(n, 0.0f), PI(n, 0.0f), PD(n, 0.0f);
    double so = 0.0, se = 0.0, ss = 0.0;
    for (int i = L1; i >= 0; --i)
        for (int j = L2; j >= 0; --j) {
            if (i == 0 && j == 0) continue;
            const size_t c = at(i, j);
            float wd[3], wi[3], wx[3];
            softmax(M[c], I[c], D[c], wd);
            softmax(M[c] + P.go2, I[c] + P.ge2, D[c] + P.go2, wi);
            softmax(M[c] + P.go2, I[c] + P.go2, D[c] + P.ge2, wx);
            const float Pd = (i < L1 && j < L2 ? PM[at(i + 1, j + 1)] : 0.0f) +
                             (i == L1 && j == L2 ? 1.0f : 0.0f);
            const float Pi = (i < L1 && j >= 1) ? PI[at(i + 1, j)] : 0.0f;  // DP cells only
            const float Px = (i >= 1 && j < L2) ? PD[at(i, j + 1)] : 0.0f;
            PM[c] = wd[0] * Pd + wi[0] * Pi + wx[0] * Px;
            PI[c] = wd[1] * Pd + wi[1] * Pi + wx[1] * Px;
            PD[c] = wd[2] * Pd + wi[2] * Pi + wx[2] * Px;
            so += (double)Pi * (wi[0] + wi[2]) + (double)Px * (wx[0] + wx[1]);
            se += (double)Pi * wi[1] + (double)Px * wx[2];
            if (j == 0) { so += PI[c]; se += (double)(i - 1) * PI[c]; }  // I(i, 0)
            if (i == 0) { so += PD[c]; se += (double)(j - 1) * PD[c]; }  // D(0, j)
            if (i >= 1 && j >= 1) {
                const int64_t q = (int64_t)(i - 1) * r.ld + j - 1;
                P.map[r.off + q] = PM[c];
                ss += (double)PM[c] * s[q];
            }
        }
    if (P.g0) P.g0[b] = (float)(P.gap_sum ? so + se : so);
    if (P.g1) P.g1[b] = (float)se;
    if (P.g2) P.g2[b] = (float)((S - (ss + (double)P.go * so + (double)P.ge * se)) / P.T);
    zero_pad(P, r, P.map);
}

// --------------------------------------------------------------------------
// Saigo-Vert, affine gap, one pair: src/sv_affine/kernels_cpu.cpp:96-356.
// Full M, I, D tables in log2 units, S = LSE(0, {M}); the pull-form adjoint
// seeds every M state with its terminal weight 2^(M - S).
// --------------------------------------------------------------------------

void sv_pair(const Problem& P, int b, bool want_bwd) {
    const PairRef r = pair_ref(P, b);
    const int L1 = r.L1, L2 = r.L2;
    if (L1 == 0 || L2 == 0) {  // the empty alignment only
        P.value[b] = 0.0f;
        if (P.g0) P.g0[b] = 0.0f;
        if (P.g1) P.g1[b] = 0.0f;
        if (P.g2) P.g2[b] = 0.0f;
        if (want_bwd) zero_pad(P, r, P.map);
        return;
    }
    const float* s = P.scores + r.off;
    const int C = L2 + 1;
    const size_t n = (size_t)(L1 + 1) * C;
    auto at = [C](int i, int j) { return (size_t)i * C + j; };
    auto l2 = [](float a, float c) {
        float hi = std::max(a, c), lo = std::min(a, c);
        return hi + lg2(1.0f + ex2(lo - hi));
    };
    std::vector<float> M(n, kNinf), I(n, kNinf), D(n, kNinf), A(n, kNinf);
    double m = 0.0;  // the empty alignment
    for (int i = 1; i <= L1; ++i)
        for (int j = 1; j <= L2; ++j) {
            const size_t c = at(i, j), dg = at(i - 1, j - 1), up = at(i - 1, j), lf = at(i, j - 1);
            M[c] = s[(int64_t)(i - 1) * r.ld + j - 1] * P.kap + l2(A[dg], 0.0f);
            I[c] = l2(M[up] + P.go2, I[up] + P.ge2);
            D[c] = l2(l2(M[lf], I[lf]) + P.go2, D[lf] + P.ge2);
            A[c] = l2(l2(M[c], I[c]), D[c]);
            m = std::max(m, (double)M[c]);
        }
    double z = std::exp2(-m);
    for (int i = 1; i <= L1; ++i)
        for (int j = 1; j <= L2; ++j) z += std::exp2((double)M[at(i, j)] - m);
    const double S2 = m + std::log2(z);
    const double S = S2 / P.kap;
    P.value[b] = (float)S;
    if (!want_bwd) return;
    auto softmax = [](float a, float c, float d, float* w) {
        float hi = std::max(std::max(a, c), d);
        float ea = ex2(a - hi), ec = ex2(c - hi), ed = ex2(d - hi), q = 1.0f / (ea + ec + ed);
        w[0] = ea * q; w[1] = ec * q; w[2] = ed * q;
    };
    std::vector<float> PM(n, 0.0f), PI(n, 0.0f), PD(n, 0.0f);
    double so = 0.0, se = 0.0, ss = 0.0;
    for (int i = L1; i >= 1; --i)
        for (int j = L2; j >= 1; --j) {
            const size_t c = at(i, j);
            float wd[3], wi[3], wx[3];
            softmax(M[c], I[c], D[c], wd);
            softmax(M[c] + P.go2, I[c] + P.ge2, kNinf, wi);
            softmax(M[c] + P.go2, I[c] + P.go2, D[c] + P.ge2, wx);
            const float sg = 1.0f / (1.0f + ex2(-A[c]));  // continuing share of M(i+1, j+1)
            const float Pd = (i < L1 && j < L2) ? PM[at(i + 1, j + 1)] * sg : 0.0f;
            const float Pi = i < L1 ? PI[at(i + 1, j)] : 0.0f;
            const float Px = j < L2 ? PD[at(i, j + 1)] : 0.0f;
            const float seed = (float)std::exp2(std::min((double)M[c] - S2, 0.0));
            PM[c] = wd[0] * Pd + wi[0] * Pi + wx[0] * Px + seed;
            PI[c] = wd[1] * Pd + wi[1] * Pi + wx[1] * Px;
            PD[c] = wd[2] * Pd + wx[2] * Px;
            so += (double)Pi * wi[0] + (double)Px * (wx[0] + wx[1]);
            se += (double)Pi * wi[1] + (double)Px * wx[2];
            const int64_t q = (int64_t)(i - 1) * r.ld + j - 1;
            P.map[r.off + q] = PM[c];
            ss += (double)PM[c] * s[q];
        }
    if (P.g0) P.g0[b] = (float)(P.gap_sum ? so + se : so);
    if (P.g1) P.g1[b] = (float)se;
    if (P.g2) P.g2[b] = (float)((S - (ss + (double)P.go * so + (double)P.ge * se)) / P.T);
    zero_pad(P, r, P.map);
}

// --------------------------------------------------------------------------
// Needleman-Wunsch affine, forward-over-reverse along a direction (v on the
// scores, dgo, dge, dT on the parameters), one pair.  All values in log2
// units x~ = x log2(e) / T, so the temperature enters only through the
// inputs s~ = s k, g~ = g k (k' = -k dT / T) and every LSE is a plain base-2
// LSE: the forward tangent of a LSE is the weighted sum of its arguments'
// tangents; the reverse carries P and P' with w' = ln2 w (z' - sum w z') for
// each base-2 softmax.  Outputs: map tangent P_M', and with `contract` the
// tangents of (opens, extends, dS/dT) along a score direction.
// --------------------------------------------------------------------------

void nw_tangent_pair(const Problem& P, int b) {
    const PairRef r = pair_ref(P, b);
    const int L1 = r.L1, L2 = r.L2;
    if (L1 == 0 || L2 == 0) {  // closed form: no score enters, the map is zero
        const int n = L1 + L2;
        if (P.value) P.value[b] = n == 0 ? 0.0f : P.go + (float)(n - 1) * P.ge;
        if (P.map) zero_pad(P, r, P.map);
        if (P.map_t) zero_pad(P, r, P.map_t);
        if (P.contract) for (int c = 0; c < 3; ++c) P.contract[3 * b + c] = 0.0f;
        return;
    }
    const float* s = P.scores + r.off;
    const float* v = P.dir ? P.dir + r.off : nullptr;
    const float k = P.kap, kd = -P.kap * P.dT / P.T;  // k'
    const float go2 = P.go2, ge2 = P.ge2;
    const float dgo2 = P.dgo * k + P.go * kd, dge2 = P.dge * k + P.ge * kd;
    const int C = L2 + 1;
    const size_t n = (size_t)(L1 + 1) * C;
    auto at = [C](int i, int j) { return (size_t)i * C + j; };
    // base-2 softmax of three values
    auto sm3 = [](float a, float c, float d, float* w) {
        float hi = std::max(std::max(a, c), d);
        float ea = ex2(a - hi), ec = ex2(c - hi), ed = ex2(d - hi), q = 1.0f / (ea + ec + ed);
        w[0] = ea * q; w[1] = ec * q; w[2] = ed * q;
    };
    auto l3 = [&](float a, float c, float d, float da, float dc, float dd, float& dy) {
        float w[3];
        sm3(a, c, d, w);
        dy = w[0] * da + w[1] * dc + w[2] * dd;
        float hi = std::max(std::max(a, c), d);
        return hi + lg2(ex2(a - hi) + ex2(c - hi) + ex2(d - hi));
    };
    std::vector<float> M(n, kNinf), I(n, kNinf), D(n, kNinf), dM(n, 0.0f), dI(n, 0.0f), dD(n, 0.0f);
    M[0] = 0.0f;
    for (int i = 1; i <= L1; ++i) {
        I[at(i, 0)] = go2 + (float)(i - 1) * ge2;
        dI[at(i, 0)] = dgo2 + (float)(i - 1) * dge2;
    }
    for (int j = 1; j <= L2; ++j) {
        D[at(0, j)] = go2 + (float)(j - 1) * ge2;
        dD[at(0, j)] = dgo2 + (float)(j - 1) * dge2;
    }
    for (int i = 1; i <= L1; ++i)
        for (int j = 1; j <= L2; ++j) {
            const size_t c = at(i, j), dg = at(i - 1, j - 1), up = at(i - 1, j), lf = at(i, j - 1);
            const int64_t q = (int64_t)(i - 1) * r.ld + j - 1;
            float dA;
            const float A = l3(M[dg], I[dg], D[dg], dM[dg], dI[dg], dD[dg], dA);
            M[c] = s[q] * k + A;
            dM[c] = (v ? v[q] * k : 0.0f) + s[q] * kd + dA;
            float di, dd;
            I[c] = l3(M[up] + go2, I[up] + ge2, D[up] + go2, dM[up] + dgo2, dI[up] + dge2,
                      dD[up] + dgo2, di);
            D[c] = l3(M[lf] + go2, I[lf] + go2, D[lf] + ge2, dM[lf] + dgo2, dI[lf] + dgo2,
                      dD[lf] + dge2, dd);
            dI[c] = di;
            dD[c] = dd;
        }
    const size_t f = at(L1, L2);
    float dS2;
    const float S2 = l3(M[f], I[f], D[f], dM[f], dI[f], dD[f], dS2);
    if (P.value) P.value[b] = S2 / k;
    std::vector<float> PM(n, 0.0f), PI(n, 0.0f), PD(n, 0.0f), tM(n, 0.0f), tI(n, 0.0f), tD(n, 0.0f);
    double so = 0.0, se = 0.0, ss = 0.0;
    const float ln2 = kLn2;
    // weights of a base-2 softmax over z and their tangents along dz
    auto wt = [&](const float* z, const float* dz, float* w, float* dw) {
        sm3(z[0], z[1], z[2], w);
        const float m = w[0] * dz[0] + w[1] * dz[1] + w[2] * dz[2];
        for (int t = 0; t < 3; ++t) dw[t] = ln2 * w[t] * (dz[t] - m);
    };
    for (int i = L1; i >= 0; --i)
        for (int j = L2; j >= 0; --j) {
            if (i == 0 && j == 0) continue;
            const size_t c = at(i, j);
            // zero tangents for the minus-infinity states
            const float dm = M[c] > kNinf ? dM[c] : 0.0f, dii = I[c] > kNinf ? dI[c] : 0.0f,
                        ddd = D[c] > kNinf ? dD[c] : 0.0f;
            float zd[3] = {M[c], I[c], D[c]}, dzd[3] = {dm, dii, ddd};
            float zi[3] = {M[c] + go2, I[c] + ge2, D[c] + go2}, dzi[3] = {dm + dgo2, dii + dge2, ddd + dgo2};
            float zx[3] = {M[c] + go2, I[c] + go2, D[c] + ge2}, dzx[3] = {dm + dgo2, dii + dgo2, ddd + dge2};
            float wd[3], wi[3], wx[3], dwd[3], dwi[3], dwx[3];
            wt(zd, dzd, wd, dwd);
            wt(zi, dzi, wi, dwi);
            wt(zx, dzx, wx, dwx);
            const bool dgn = i < L1 && j < L2, iin = i < L1 && j >= 1, xin = i >= 1 && j < L2;
            const float Pd = (dgn ? PM[at(i + 1, j + 1)] : 0.0f) + (i == L1 && j == L2 ? 1.0f : 0.0f);
            const float Td = dgn ? tM[at(i + 1, j + 1)] : 0.0f;
            const float Pi = iin ? PI[at(i + 1, j)] : 0.0f, Ti = iin ? tI[at(i + 1, j)] : 0.0f;
            const float Px = xin ? PD[at(i, j + 1)] : 0.0f, Tx = xin ? tD[at(i, j + 1)] : 0.0f;
            float* Pc[3] = {&PM[c], &PI[c], &PD[c]};
            float* Tc[3] = {&tM[c], &tI[c], &tD[c]};
            for (int t = 0; t < 3; ++t) {
                *Pc[t] = wd[t] * Pd + wi[t] * Pi + wx[t] * Px;
                *Tc[t] = dwd[t] * Pd + wd[t] * Td + dwi[t] * Pi + wi[t] * Ti + dwx[t] * Px + wx[t] * Tx;
            }
            // tangents of the open / extend flows
            so += (double)Ti * (wi[0] + wi[2]) + (double)Pi * (dwi[0] + dwi[2]) +
                  (double)Tx * (wx[0] + wx[1]) + (double)Px * (dwx[0] + dwx[1]);
            se += (double)Ti * wi[1] + (double)Pi * dwi[1] + (double)Tx * wx[2] + (double)Px * dwx[2];
            if (j == 0) { so += tI[c]; se += (double)(i - 1) * tI[c]; }  // I(i, 0)
            if (i == 0) { so += tD[c]; se += (double)(j - 1) * tD[c]; }  // D(0, j)
            if (i >= 1 && j >= 1) {
                const int64_t q = (int64_t)(i - 1) * r.ld + j - 1;
                if (P.map) P.map[r.off + q] = PM[c];
                if (P.map_t) P.map_t[r.off + q] = tM[c];
                ss += (double)tM[c] * s[q];
            }
        }
    if (P.contract) {
        P.contract[3 * b + 0] = (float)so;
        P.contract[3 * b + 1] = (float)se;
        P.contract[3 * b + 2] = (float)(-(ss + (double)P.go * so + (double)P.ge * se) / P.T);
    }
    if (P.map) zero_pad(P, r, P.map);
    if (P.map_t) zero_pad(P, r, P.map_t);
}

// --------------------------------------------------------------------------
// Saigo-Vert affine, forward-over-reverse along a direction, one pair (log2
// units as nw_tangent_pair): the restart split sig(A) has the tangent
// ln2 sig (1 - sig) A', the terminal seeds 2^(M - S) the tangent
// ln2 2^(M - S) (M' - S') with S' = sum 2^(M - S) M'.
// --------------------------------------------------------------------------

void sv_tangent_pair(const Problem& P, int b) {
    const PairRef r = pair_ref(P, b);
    const int L1 = r.L1, L2 = r.L2;
    if (L1 == 0 || L2 == 0) {
        if (P.value) P.value[b] = 0.0f;
        if (P.map) zero_pad(P, r, P.map);
        if (P.map_t) zero_pad(P, r, P.map_t);
        if (P.contract) for (int c = 0; c < 3; ++c) P.contract[3 * b + c] = 0.0f;
        return;
    }
    const float* s = P.scores + r.off;
    const float* v = P.dir ? P.dir + r.off : nullptr;
    const float k = P.kap, kd = -P.kap * P.dT / P.T;
    const float go2 = P.go2, ge2 = P.ge2;
    const float dgo2 = P.dgo * k + P.go * kd, dge2 = P.dge * k + P.ge * kd;
    const int C = L2 + 1;
    const size_t n = (size_t)(L1 + 1) * C;
    auto at = [C](int i, int j) { return (size_t)i * C + j; };
    auto sm3 = [](float a, float c, float d, float* w) {
        float hi = std::max(std::max(a, c), d);
        float ea = ex2(a - hi), ec = ex2(c - hi), ed = ex2(d - hi), q = 1.0f / (ea + ec + ed);
        w[0] = ea * q; w[1] = ec * q; w[2] = ed * q;
    };
    auto l3 = [&](float a, float c, float d, float da, float dc, float dd, float& dy) {
        float w[3];
        sm3(a, c, d, w);
        dy = w[0] * da + w[1] * dc + w[2] * dd;
        float hi = std::max(std::max(a, c), d);
        return hi + lg2(ex2(a - hi) + ex2(c - hi) + ex2(d - hi));
    };
    std::vector<float> M(n, kNinf), I(n, kNinf), D(n, kNinf), A(n, kNinf);
    std::vector<float> dM(n, 0.0f), dI(n, 0.0f), dD(n, 0.0f), dA(n, 0.0f);
    double m = 0.0;
    for (int i = 1; i <= L1; ++i)
        for (int j = 1; j <= L2; ++j) {
            const size_t c = at(i, j), dg = at(i - 1, j - 1), up = at(i - 1, j), lf = at(i, j - 1);
            const int64_t q = (int64_t)(i - 1) * r.ld + j - 1;
            float dsp;  // LSE(A_diag, 0) and its tangent
            const float sp = l3(A[dg], 0.0f, kNinf, dA[dg], 0.0f, 0.0f, dsp);
            M[c] = s[q] * k + sp;
            dM[c] = (v ? v[q] * k : 0.0f) + s[q] * kd + dsp;
            float di, dd, da;
            I[c] = l3(M[up] + go2, I[up] + ge2, kNinf, dM[up] + dgo2, dI[up] + dge2, 0.0f, di);
            D[c] = l3(M[lf] + go2, I[lf] + go2, D[lf] + ge2, dM[lf] + dgo2, dI[lf] + dgo2,
                      dD[lf] + dge2, dd);
            dI[c] = di;
            dD[c] = dd;
            A[c] = l3(M[c], I[c], D[c], dM[c], dI[c], dD[c], da);
            dA[c] = da;
            m = std::max(m, (double)M[c]);
        }
    double z = std::exp2(-m), zt = 0.0;
    for (int i = 1; i <= L1; ++i)
        for (int j = 1; j <= L2; ++j) {
            const double w = std::exp2((double)M[at(i, j)] - m);
            z += w;
            zt += w * dM[at(i, j)];
        }
    const double S2 = m + std::log2(z), dS2 = zt / z;
    if (P.value) P.value[b] = (float)(S2 / k);
    const float ln2 = kLn2;
    auto wt = [&](const float* zz, const float* dz, float* w, float* dw) {
        sm3(zz[0], zz[1], zz[2], w);
        const float mm = w[0] * dz[0] + w[1] * dz[1] + w[2] * dz[2];
        for (int t = 0; t < 3; ++t) dw[t] = ln2 * w[t] * (dz[t] - mm);
    };
    std::vector<float> PM(n, 0.0f), PI(n, 0.0f), PD(n, 0.0f), tM(n, 0.0f), tI(n, 0.0f), tD(n, 0.0f);
    double so = 0.0, se = 0.0, ss = 0.0;
    for (int i = L1; i >= 1; --i)
        for (int j = L2; j >= 1; --j) {
            const size_t c = at(i, j);
            float zd[3] = {M[c], I[c], D[c]}, dzd[3] = {dM[c], dI[c], dD[c]};
            float zi[3] = {M[c] + go2, I[c] + ge2, kNinf}, dzi[3] = {dM[c] + dgo2, dI[c] + dge2, 0.0f};
            float zx[3] = {M[c] + go2, I[c] + go2, D[c] + ge2};
            float dzx[3] = {dM[c] + dgo2, dI[c] + dgo2, dD[c] + dge2};
            float wd[3], wi[3], wx[3], dwd[3], dwi[3], dwx[3];
            wt(zd, dzd, wd, dwd);
            wt(zi, dzi, wi, dwi);
            wt(zx, dzx, wx, dwx);
            const float sg = 1.0f / (1.0f + ex2(-A[c])), dsg = ln2 * sg * (1.0f - sg) * dA[c];
            const bool dgn = i < L1 && j < L2;
            const float Pd = dgn ? PM[at(i + 1, j + 1)] : 0.0f, Td = dgn ? tM[at(i + 1, j + 1)] : 0.0f;
            const float Pi = i < L1 ? PI[at(i + 1, j)] : 0.0f, Ti = i < L1 ? tI[at(i + 1, j)] : 0.0f;
            const float Px = j < L2 ? PD[at(i, j + 1)] : 0.0f, Tx = j < L2 ? tD[at(i, j + 1)] : 0.0f;
            const double e = std::exp2(std::min((double)M[c] - S2, 0.0));
            const float seed = (float)e, dseed = (float)(e * kLn2 * ((double)dM[c] - dS2));
            float* Pc[3] = {&PM[c], &PI[c], &PD[c]};
            float* Tc[3] = {&tM[c], &tI[c], &tD[c]};
            for (int t = 0; t < 3; ++t) {
                *Pc[t] = wd[t] * sg * Pd + wi[t] * Pi + wx[t] * Px + (t == 0 ? seed : 0.0f);
                *Tc[t] = (dwd[t] * sg + wd[t] * dsg) * Pd + wd[t] * sg * Td + dwi[t] * Pi +
                         wi[t] * Ti + dwx[t] * Px + wx[t] * Tx + (t == 0 ? dseed : 0.0f);
            }
            so += (double)Ti * wi[0] + (double)Pi * dwi[0] + (double)Tx * (wx[0] + wx[1]) +
                  (double)Px * (dwx[0] + dwx[1]);
            se += (double)Ti * wi[1] + (double)Pi * dwi[1] + (double)Tx * wx[2] + (double)Px * dwx[2];
            const int64_t q = (int64_t)(i - 1) * r.ld + j - 1;
            if (P.map) P.map[r.off + q] = PM[c];
            if (P.map_t) P.map_t[r.off + q] = tM[c];
            ss += (double)tM[c] * s[q];
        }
    if (P.contract) {
        P.contract[3 * b + 0] = (float)so;
        P.contract[3 * b + 1] = (float)se;
        P.contract[3 * b + 2] = (float)(-(ss + (double)P.go * so + (double)P.ge * se) / P.T);
    }
    if (P.map) zero_pad(P, r, P.map);
    if (P.map_t) zero_pad(P, r, P.map_t);
}

template <typename F>
void parallel_pairs(int B, int threads, F&& fn) {
    int nt = threads > 0 ? threads : (int)std::thread::hardware_concurrency();
    if (nt < 1) nt = 1;
    if (nt > B) nt = B;
    if (nt <= 1) {
        for (int b = 0; b < B; ++b) fn(b);
        return;
    }
    std::atomic<int> next(0);
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t)
        pool.emplace_back([&]() {
            for (int b = next++; b < B; b = next++) fn(b);
        });
    for (auto& th : pool) th.join();
}

ohm_status check(const Problem& P) {
    if (P.B < 0 || P.L1max < 0 || P.L2max < 0) return cpu_fail(OHM_EINVAL, "negative batch size or length");
    if (!(P.T > 0.0f) || !std::isfinite(P.T))
        return cpu_fail(OHM_EINVAL, "temperature must be finite and strictly positive");
    // a batch with an empty dimension has no cells: its scores may be null
    if (P.B > 0 && P.L1max > 0 && P.L2max > 0 && !P.scores)
        return cpu_fail(OHM_EINVAL, "scores is null");
    if (P.offsets && !P.lengths) return cpu_fail(OHM_EINVAL, "the packed layout needs lengths");
    for (int b = 0; P.lengths && b < P.B; ++b) {
        int l1 = P.lengths[2 * b], l2 = P.lengths[2 * b + 1];
        if (l1 < 0 || l2 < 0 || (!P.offsets && (l1 > P.L1max || l2 > P.L2max)))
            return cpu_fail(OHM_EINVAL, "lengths out of range");
    }
    return OHM_OK;
}

Problem make(const float* scores, const int32_t* lengths, const int64_t* offsets, int B, int L1max,
             int L2max, float go, float ge, float T) {
    Problem P;
    std::memset(&P, 0, sizeof(P));
    P.scores = scores;
    P.lengths = lengths;
    P.offsets = offsets;
    P.B = B;
    P.L1max = L1max;
    P.L2max = L2max;
    P.go = go;
    P.ge = ge;
    P.T = T;
    P.kap = kLog2e / T;
    P.go2 = go * P.kap;
    P.ge2 = ge * P.kap;
    return P;
}

}  // namespace

extern "C" ohm_status ohm_sw_affine_forward_cpu(const float* scores, const int32_t* lengths,
                                                const int64_t* offsets, int32_t B, int32_t L1max,
                                                int32_t L2max, float gap_open, float gap_ext,
                                                float temp, float* value, float* map,
                                                float* grad_open, float* grad_ext,
                                                float* grad_temp, int32_t threads) {
    Problem P = make(scores, lengths, offsets, B, L1max, L2max, gap_open, gap_ext, temp);
    P.value = value; P.map = map; P.g0 = grad_open; P.g1 = grad_ext; P.g2 = grad_temp;
    ohm_status st = check(P);
    if (st != OHM_OK) return st;
    const bool bwd = map || grad_open || grad_ext || grad_temp;
    parallel_pairs(B, threads, [&](int b) { affine_pair<false>(P, b, bwd); });
    return OHM_OK;
}

extern "C" ohm_status ohm_nw_affine_forward_cpu(const float* scores, const int32_t* lengths,
                                                const int64_t* offsets, int32_t B, int32_t L1max,
                                                int32_t L2max, float gap_open, float gap_ext,
                                                float temp, float* value, float* map,
                                                float* grad_open, float* grad_ext,
                                                float* grad_temp, int32_t threads) {
    Problem P = make(scores, lengths, offsets, B, L1max, L2max, gap_open, gap_ext, temp);
    P.value = value; P.map = map; P.g0 = grad_open; P.g1 = grad_ext; P.g2 = grad_temp;
    ohm_status st = check(P);
    if (st != OHM_OK) return st;
    if (!std::isfinite(gap_open) || !std::isfinite(gap_ext))
        return cpu_fail(OHM_EINVAL, "gap scores must be finite");
    const bool bwd = map || grad_open || grad_ext || grad_temp;
    if (bwd && !map && L1max > 0 && L2max > 0)
        return cpu_fail(OHM_EINVAL, "the gradients need the map buffer");
    parallel_pairs(B, threads, [&](int b) { nw_pair(P, b, bwd); });
    return OHM_OK;
}

extern "C" ohm_status ohm_sv_affine_forward_cpu(const float* scores, const int32_t* lengths,
                                                const int64_t* offsets, int32_t B, int32_t L1max,
                                                int32_t L2max, float gap_open, float gap_ext,
                                                float temp, float* value, float* map,
                                                float* grad_open, float* grad_ext,
                                                float* grad_temp, int32_t threads) {
    Problem P = make(scores, lengths, offsets, B, L1max, L2max, gap_open, gap_ext, temp);
    P.value = value; P.map = map; P.g0 = grad_open; P.g1 = grad_ext; P.g2 = grad_temp;
    ohm_status st = check(P);
    if (st != OHM_OK) return st;
    if (!std::isfinite(gap_open) || !std::isfinite(gap_ext))
        return cpu_fail(OHM_EINVAL, "gap scores must be finite");
    const bool bwd = map || grad_open || grad_ext || grad_temp;
    if (bwd && !map && L1max > 0 && L2max > 0)
        return cpu_fail(OHM_EINVAL, "the gradients need the map buffer");
    parallel_pairs(B, threads, [&](int b) { sv_pair(P, b, bwd); });
    return OHM_OK;
}

extern "C" ohm_status ohm_nw_forward_cpu(const float* scores, const int32_t* lengths,
                                         const int64_t* offsets, int32_t B, int32_t L1max,
                                         int32_t L2max, float gap, float temp, float* value,
                                         float* map, float* grad_gap, float* grad_temp,
                                         int32_t threads) {
    Problem P = make(scores, lengths, offsets, B, L1max, L2max, gap, gap, temp);
    P.value = value; P.map = map; P.g0 = grad_gap; P.g2 = grad_temp; P.gap_sum = true;
    ohm_status st = check(P);
    if (st != OHM_OK) return st;
    if (!std::isfinite(gap)) return cpu_fail(OHM_EINVAL, "gap scores must be finite");
    const bool bwd = map || grad_gap || grad_temp;
    if (bwd && !map && L1max > 0 && L2max > 0)
        return cpu_fail(OHM_EINVAL, "the gradients need the map buffer");
    parallel_pairs(B, threads, [&](int b) { nw_pair(P, b, bwd); });
    return OHM_OK;
}

extern "C" ohm_status ohm_sv_forward_cpu(const float* scores, const int32_t* lengths,
                                         const int64_t* offsets, int32_t B, int32_t L1max,
                                         int32_t L2max, float gap, float temp, float* value,
                                         float* map, float* grad_gap, float* grad_temp,
                                         int32_t threads) {
    Problem P = make(scores, lengths, offsets, B, L1max, L2max, gap, gap, temp);
    P.value = value; P.map = map; P.g0 = grad_gap; P.g2 = grad_temp; P.gap_sum = true;
    ohm_status st = check(P);
    if (st != OHM_OK) return st;
    if (!std::isfinite(gap)) return cpu_fail(OHM_EINVAL, "gap scores must be finite");
    const bool bwd = map || grad_gap || grad_temp;
    if (bwd && !map && L1max > 0 && L2max > 0)
        return cpu_fail(OHM_EINVAL, "the gradients need the map buffer");
    parallel_pairs(B, threads, [&](int b) { sv_pair(P, b, bwd); });
    return OHM_OK;
}

extern "C" ohm_status ohm_nw_affine_tangent_cpu(const float* scores, const int32_t* lengths,
                                                const int64_t* offsets, int32_t B, int32_t L1max,
                                                int32_t L2max, float gap_open, float gap_ext,
                                                float temp, const float* direction,
                                                float d_gap_open, float d_gap_ext, float d_temp,
                                                float* map_tangent, float* contract, float* value,
                                                float* map, int32_t threads) {
    Problem P = make(scores, lengths, offsets, B, L1max, L2max, gap_open, gap_ext, temp);
    P.dir = direction; P.dgo = d_gap_open; P.dge = d_gap_ext; P.dT = d_temp;
    P.map_t = map_tangent; P.contract = contract; P.value = value; P.map = map;
    ohm_status st = check(P);
    if (st != OHM_OK) return st;
    if (contract && (d_gap_open != 0.0f || d_gap_ext != 0.0f || d_temp != 0.0f))
        return cpu_fail(OHM_EINVAL, "contract needs a pure score direction");
    parallel_pairs(B, threads, [&](int b) { nw_tangent_pair(P, b); });
    return OHM_OK;
}

extern "C" ohm_status ohm_sv_affine_tangent_cpu(const float* scores, const int32_t* lengths,
                                                const int64_t* offsets, int32_t B, int32_t L1max,
                                                int32_t L2max, float gap_open, float gap_ext,
                                                float temp, const float* direction,
                                                float d_gap_open, float d_gap_ext, float d_temp,
                                                float* map_tangent, float* contract, float* value,
                                                float* map, int32_t threads) {
    Problem P = make(scores, lengths, offsets, B, L1max, L2max, gap_open, gap_ext, temp);
    P.dir = direction; P.dgo = d_gap_open; P.dge = d_gap_ext; P.dT = d_temp;
    P.map_t = map_tangent; P.contract = contract; P.value = value; P.map = map;
    ohm_status st = check(P);
    if (st != OHM_OK) return st;
    if (contract && (d_gap_open != 0.0f || d_gap_ext != 0.0f || d_temp != 0.0f))
        return cpu_fail(OHM_EINVAL, "contract needs a pure score direction");
    parallel_pairs(B, threads, [&](int b) { sv_tangent_pair(P, b); });
    return OHM_OK;
}

extern "C" ohm_status ohm_sw_affine_tangent_cpu(const float* scores, const int32_t* lengths,
                                                const int64_t* offsets, int32_t B, int32_t L1max,
                                                int32_t L2max, float gap_open, float gap_ext,
                                                float temp, const float* direction,
                                                float d_gap_open, float d_gap_ext, float d_temp,
                                                float* map_tangent, float* contract, float* value,
                                                float* map, int32_t threads) {
    Problem P = make(scores, lengths, offsets, B, L1max, L2max, gap_open, gap_ext, temp);
    P.dir = direction; P.dgo = d_gap_open; P.dge = d_gap_ext; P.dT = d_temp;
    P.map_t = map_tangent; P.contract = contract; P.value = value; P.map = map;
    ohm_status st = check(P);
    if (st != OHM_OK) return st;
    if (contract && (d_gap_open != 0.0f || d_gap_ext != 0.0f || d_temp != 0.0f))
        return cpu_fail(OHM_EINVAL, "contract needs a pure score direction");
    parallel_pairs(B, threads, [&](int b) { affine_pair<true>(P, b, true); });
    return OHM_OK;
}

extern "C" ohm_status ohm_sw_forward_cpu(const float* scores, const int32_t* lengths,
                                         const int64_t* offsets, int32_t B, int32_t L1max,
                                         int32_t L2max, float gap, float temp, float* value,
                                         float* map, float* grad_gap, float* grad_temp,
                                         int32_t threads) {
    Problem P = make(scores, lengths, offsets, B, L1max, L2max, gap, 0.0f, temp);
    P.value = value; P.map = map; P.g0 = grad_gap; P.g1 = grad_temp;
    ohm_status st = check(P);
    if (st != OHM_OK) return st;
    const bool bwd = map || grad_gap || grad_temp;
    parallel_pairs(B, threads, [&](int b) { linear_pair<false>(P, b, bwd); });
    return OHM_OK;
}

extern "C" ohm_status ohm_sw_tangent_cpu(const float* scores, const int32_t* lengths,
                                         const int64_t* offsets, int32_t B, int32_t L1max,
                                         int32_t L2max, float gap, float temp,
                                         const float* direction, float d_gap, float d_temp,
                                         float* map_tangent, float* contract, float* value,
                                         float* map, int32_t threads) {
    Problem P = make(scores, lengths, offsets, B, L1max, L2max, gap, 0.0f, temp);
    P.dir = direction; P.dgo = d_gap; P.dT = d_temp;
    P.map_t = map_tangent; P.contract = contract; P.value = value; P.map = map;
    ohm_status st = check(P);
    if (st != OHM_OK) return st;
    if (contract && (d_gap != 0.0f || d_temp != 0.0f))
        return cpu_fail(OHM_EINVAL, "contract needs a pure score direction");
    parallel_pairs(B, threads, [&](int b) { linear_pair<true>(P, b, true); });
    return OHM_OK;
}

// One error slot per thread, shared by the CPU and GPU entry points.
namespace ohm {
static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
}  // namespace ohm

extern "C" const char* ohm_last_error(void) { return ohm::g_last_error.c_str(); }
extern "C" const char* ohm_version(void) { return "0.1.0"; }
