// Host implementation of include/ohm_embed.h: the same products as the
// gfx950 kernels (BF16 inputs, FP32 accumulation in index order), one thread
// per pair block.  dQ / dK use the FP32 map as is.

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "ohm_embed.h"

namespace ohm {
void set_error(const std::string& msg);
}

namespace {

float bf16_to_float(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

uint16_t float_to_bf16(float x) {  // round to nearest even (finite inputs)
    uint32_t u;
    std::memcpy(&u, &x, 4);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

ohm_status fail(ohm_status st, const char* msg) {
    ohm::set_error(msg);
    return st;
}

template <typename F>
void each_pair(int B, F&& fn) {
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t)
        pool.emplace_back([&, t] {
            for (int b = t; b < B; b += nt) fn(b);
        });
    for (auto& th : pool) th.join();
}

}  // namespace

extern "C" ohm_status ohm_embed_scores_cpu(const void* Q, const void* K, const int32_t* lengths,
                                           int32_t B, int32_t Lq, int32_t Lk, int32_t D,
                                           int64_t* offsets, float* scores) {
    if (B < 0 || Lq < 0 || Lk < 0 || D <= 0) return fail(OHM_EINVAL, "invalid dimensions");
    if (B == 0) return OHM_OK;
    if (!lengths || !offsets) return fail(OHM_EINVAL, "lengths and offsets must be non-null");
    int64_t run = 0;
    for (int b = 0; b < B; ++b) {
        offsets[b] = run;
        run += (int64_t)clampi(lengths[2 * b], 0, Lq) * clampi(lengths[2 * b + 1], 0, Lk);
    }
    if (run == 0) return OHM_OK;
    if (!Q || !K || !scores) return fail(OHM_EINVAL, "Q, K and scores must be non-null");
    const uint16_t* q = static_cast<const uint16_t*>(Q);
    const uint16_t* k = static_cast<const uint16_t*>(K);
    each_pair(B, [&](int b) {
        const int M = clampi(lengths[2 * b], 0, Lq), N = clampi(lengths[2 * b + 1], 0, Lk);
        std::vector<float> qi(D);
        for (int i = 0; i < M; ++i) {
            const uint16_t* qr = q + ((size_t)b * Lq + i) * D;
            for (int d = 0; d < D; ++d) qi[d] = bf16_to_float(qr[d]);
            for (int j = 0; j < N; ++j) {
                const uint16_t* kr = k + ((size_t)b * Lk + j) * D;
                float acc = 0.0f;
                for (int d = 0; d < D; ++d) acc += qi[d] * bf16_to_float(kr[d]);
                scores[offsets[b] + (int64_t)i * N + j] = acc;
            }
        }
    });
    return OHM_OK;
}

extern "C" ohm_status ohm_embed_grads_cpu(const void* Q, const void* K, const int32_t* lengths,
                                          const int64_t* offsets, int32_t B, int32_t Lq,
                                          int32_t Lk, int32_t D, const float* map,
                                          const float* grad_value, void* dQ, void* dK) {
    if (B < 0 || Lq < 0 || Lk < 0 || D <= 0) return fail(OHM_EINVAL, "invalid dimensions");
    if (B == 0) return OHM_OK;
    if (!lengths || !offsets || !grad_value) return fail(OHM_EINVAL, "lengths, offsets and grad_value must be non-null");
    if ((Lq > 0 && !dQ) || (Lk > 0 && !dK)) return fail(OHM_EINVAL, "dQ and dK must be non-null");
    const uint16_t* q = static_cast<const uint16_t*>(Q);
    const uint16_t* k = static_cast<const uint16_t*>(K);
    uint16_t* gq = static_cast<uint16_t*>(dQ);
    uint16_t* gk = static_cast<uint16_t*>(dK);
    each_pair(B, [&](int b) {
        const int M = clampi(lengths[2 * b], 0, Lq), N = clampi(lengths[2 * b + 1], 0, Lk);
        const float g = grad_value[b];
        const float* P = map ? map + offsets[b] : nullptr;
        std::vector<float> acc(D);
        for (int i = 0; i < Lq; ++i) {
            std::fill(acc.begin(), acc.end(), 0.0f);
            if (i < M)
                for (int j = 0; j < N; ++j) {
                    const float p = P[(int64_t)i * N + j];
                    const uint16_t* kr = k + ((size_t)b * Lk + j) * D;
                    for (int d = 0; d < D; ++d) acc[d] += p * bf16_to_float(kr[d]);
                }
            uint16_t* out = gq + ((size_t)b * Lq + i) * D;
            for (int d = 0; d < D; ++d) out[d] = float_to_bf16(i < M ? g * acc[d] : 0.0f);
        }
        for (int j = 0; j < Lk; ++j) {
            std::fill(acc.begin(), acc.end(), 0.0f);
            if (j < N)
                for (int i = 0; i < M; ++i) {
                    const float p = P[(int64_t)i * N + j];
                    const uint16_t* qr = q + ((size_t)b * Lq + i) * D;
                    for (int d = 0; d < D; ++d) acc[d] += p * bf16_to_float(qr[d]);
                }
            uint16_t* out = gk + ((size_t)b * Lk + j) * D;
            for (int d = 0; d < D; ++d) out[d] = float_to_bf16(j < N ? g * acc[d] : 0.0f);
        }
    });
    return OHM_OK;
}

extern "C" ohm_status ohm_embed_sw_affine_forward_cpu(const void* Q, const void* K,
                                                      const int32_t* lengths, int32_t B, int32_t Lq,
                                                      int32_t Lk, int32_t D, float gap_open,
                                                      float gap_ext, float temp, int64_t* offsets,
                                                      float* scores, float* value, float* map,
                                                      float* grad_open, float* grad_ext,
                                                      float* grad_temp, int32_t threads) {
    if (D != 32 && D != 64 && D != 128) return fail(OHM_EINVAL, "the fused embedding sweeps take D = 32, 64 or 128");
    ohm_status st = ohm_embed_scores_cpu(Q, K, lengths, B, Lq, Lk, D, offsets, scores);
    if (st != OHM_OK || B == 0) return st;
    return ohm_sw_affine_forward_cpu(scores, lengths, offsets, B, Lq, Lk, gap_open, gap_ext, temp, value,
                                     map, grad_open, grad_ext, grad_temp, threads);
}
