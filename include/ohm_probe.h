/* Issue-rate probes (libohm_probe.so): measure the vector-ALU, transcendental
 * and DPP instruction rates that bound the soft Smith-Waterman sweeps.  No
 * counterpart in the reference: its roofline constants are assumptions, these
 * kernels measure them on the device the sweeps run on. */
#ifndef OHM_PROBE_H
#define OHM_PROBE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    OHM_PROBE_EXP = 0, /* v_exp_f32 */
    OHM_PROBE_LOG = 1, /* v_log_f32 */
    OHM_PROBE_FMA = 2, /* v_fma_f32 */
    OHM_PROBE_DPP = 3, /* v_mov_b32_dpp wave_shr:1 */
    OHM_PROBE_MIX = 4  /* one v_exp_f32 and four independent v_fma_f32 */
};

/* Lane-operations one thread executes for (kind, iters); 0 on success. */
int ohm_probe_ops_per_thread(int kind, int iters, int64_t* ops);

/* Launch `blocks` workgroups of 256 threads on `stream`; each thread runs the
 * probe for `iters` iterations and writes one float to sink[thread id]
 * (sink holds blocks * 256 floats).  0 on success, 1 bad argument, 2 launch
 * error. */
int ohm_probe_launch(int kind, int blocks, int iters, float* sink, void* stream);

#ifdef __cplusplus
}
#endif

#endif
