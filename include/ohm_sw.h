/*
 * ohm_sw.h — C ABI of the MI355X (gfx950) soft Smith-Waterman hot path.
 *
 * One shared library, orihime_amd/libohm_sw.so, exports everything declared
 * here.  Signatures use plain pointers, sizes and scalars only (no torch
 * types), so the library can be bound from ctypes, cgo, JNI or N-API.
 *
 * The operations replace the reference's native layer for soft
 * Smith-Waterman (linear gap, src/sw; affine gap, src/sw_affine):
 *   - the dispatcher schemas src/sw_affine/registry.cpp:27-42 and
 *     src/sw/registry.cpp:30-44 (what orihime/_ops.py:434-449 calls), and
 *   - the extern "C" kernel launchers below them,
 *     src/sw_affine/kernels_gpu.hiph:165-265, src/sw/kernels_gpu.hiph,
 *     src/sw_affine/kernels_cpu.h:80-163, src/sw/kernels_cpu.h.
 * Each declaration names the reference primitive(s) it serves.
 *
 * Problem description, shared by every entry point:
 *   scores    FP32 pair scores.  offsets == NULL: padded [B, L1max, L2max]
 *             row-major (row stride L2max).  offsets != NULL: packed — pair b
 *             occupies lengths[b,0] x lengths[b,1] floats starting at element
 *             offsets[b] (row stride lengths[b,1]).
 *   lengths   int32 [B, 2] active lengths (L1, L2) of each pair, 0 <= L <= max.
 *             NULL means every pair uses the full (L1max, L2max) (padded only).
 *             The GPU kernels read lengths on the device and clamp them into
 *             [0, max]; they do not report an out-of-range length (host
 *             callers validate, as the reference bindings do).
 *   offsets   int64 [B] or NULL, see above.  Precondition of the packed
 *             layout (not checked on the device): for every pair,
 *             0 <= offsets[b] and offsets[b] + L1_b * L2_b <= total cells of
 *             the scores buffer (L clamped as above), and the sum of
 *             L1_b * L2_b is the total_cells given to the *_workspace query.
 *             A workspace sized for fewer cells than the lengths ask for
 *             turns the pairs that do not fit into NaN outputs.
 *   Map-shaped outputs and directions use the same layout as scores.  In the
 *   padded layout, cells outside a pair's lengths are written as 0.
 *   All GPU pointers are device pointers; *_cpu entry points take host
 *   pointers.
 *
 * GPU contract: the call only enqueues work on `stream` (a hipStream_t, NULL =
 * default stream).  It allocates nothing, never synchronises the host, and
 * reads parameters by value, so it can be captured in a HIP graph.  The
 * caller provides `workspace` of at least the size returned by the matching
 * *_workspace query.  Results are bitwise repeatable: every reduction runs in
 * a fixed order and no floating-point atomics are used.
 *
 * Errors are returned as status codes (never thrown); ohm_last_error() gives
 * a message for the last failure on the calling thread.
 */
#ifndef OHM_SW_H
#define OHM_SW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ohm_status {
    OHM_OK = 0,
    OHM_EINVAL = 1,      /* invalid argument (shape, length, temperature, pointer) */
    OHM_EWORKSPACE = 2,  /* workspace missing or smaller than required */
    OHM_ELAUNCH = 3,     /* HIP launch or runtime failure */
} ohm_status;

/* Message describing the last non-OK status returned on this thread. */
const char* ohm_last_error(void);

/* Library version, "major.minor.patch". */
const char* ohm_version(void);

/* Kernel timing hook (diagnostics; the reference has no counterpart).  While
 * set, every GPU entry point records events[k] (a hipEvent_t created by the
 * caller) on its stream after its launches: k = 0 after the planning kernel,
 * k = 1 after the forward sweep (or tangent forward sweep), k = 2 after the
 * forward sweep's log-domain launch for flagged pairs, k = 3 after the
 * reverse sweep, k = 4 after the reverse sweep's log-domain launch (marks
 * without such a launch are recorded right after the previous one).  NULL
 * entries are skipped; count 0 clears the hook (count <= 5).
 * The setting is process-wide and unsynchronised: change it only while no
 * other thread is inside the library, and never while a HIP graph is being
 * captured. */
ohm_status ohm_set_stage_events(void* const* events, int32_t count);

/* Launch-configuration hook (tuning; the reference has no counterpart, its
 * launches have fixed shapes).  waves: waves per workgroup of the per-pair
 * kernels (1, 2, 4, 8; 0 = automatic); groups_per_cu: persistent workgroups
 * per CU (1..64; 0 = automatic); teams: 0 automatic, 1 wave-scheduled team
 * kernels, 2 per-pair workgroup kernels, 3 grid kernels (the strips of every
 * pair spread over all CUs; batches of at most 64 pairs), 4 per-pair
 * workgroup kernels that run each pair's forward and reverse sweep in one
 * launch (affine sweeps; the automatic choice where the per-pair kernels
 * run, 2 keeps separate forward and reverse launches).  Results do not
 * depend on the configuration, only the speed.  Process-wide and
 * unsynchronised like ohm_set_stage_events; orihime_amd/tune.py chooses it
 * by timing the caller's batch. */
ohm_status ohm_set_launch_config(int32_t waves, int32_t groups_per_cu, int32_t teams);

/* Tile hook (tuning; no reference counterpart: the reference's adjoint keeps
 * whole DP tables).  steps: anti-diagonal steps between the checkpoints that
 * the affine forward sweep stores for the reverse sweep (4, 8 or 16; 0 =
 * automatic, 8).  Shorter strides store more and recompute less.  The
 * checkpoints are exact sweep states, so results are bitwise the same for
 * every stride.  The workspace queries use the stride set when they are
 * called: set it before sizing a workspace.  Process-wide and unsynchronised
 * like ohm_set_launch_config. */
ohm_status ohm_set_checkpoint_stride(int32_t steps);
/* The stride in effect (after resolving 0). */
int32_t ohm_checkpoint_stride(void);

/* Operation kinds for the workspace queries. */
#define OHM_OP_VALUE 0    /* value only */
#define OHM_OP_FORWARD 1  /* value + map + parameter gradients */
#define OHM_OP_TANGENT 2  /* forward-over-reverse: HVP, sensitivity, map cotangent */

/* Workspace bytes for one call.  total_cells is sum_b L1_b*L2_b for the packed
 * layout, or -1 for the padded layout (bounded by B*L1max*L2max). */
size_t ohm_sw_affine_workspace(int32_t op, int32_t B, int32_t L1max, int32_t L2max,
                               int64_t total_cells);
size_t ohm_sw_workspace(int32_t op, int32_t B, int32_t L1max, int32_t L2max,
                        int64_t total_cells);

/* ------------------------------------------------------------------------ */
/* Affine gap (three states M, I, D)                                         */
/* ------------------------------------------------------------------------ */

/* Value, marginal map and scalar-parameter gradients.
 *   value[B]                 partition value S_b (always written)
 *   map (nullable)           dS/dscores, padded-layout zeros outside lengths
 *   grad_open/ext/temp[B]    (nullable) dS/d(gap_open), dS/d(gap_ext),
 *                            dS/d(temperature) = entropy
 * When map and all three gradients are NULL only the forward sweep runs.
 * Replaces: sw_affine_forward, sw_affine_forward_t (forward part),
 * sw_affine_value_grad_params, soft_sw_affine_with_grads
 * (src/sw_affine/torch_hip.cpp:128-208, :446, :820) and the launchers
 * sw_affine_forward_hip / sw_affine_backward_hip (kernels_gpu.hiph). */
ohm_status ohm_sw_affine_forward(const float* scores, const int32_t* lengths,
                                 const int64_t* offsets, int32_t B, int32_t L1max,
                                 int32_t L2max, float gap_open, float gap_ext, float temp,
                                 float* value, float* map, float* grad_open, float* grad_ext,
                                 float* grad_temp, void* workspace, size_t workspace_bytes,
                                 void* stream);

/* Forward-over-reverse pass along one direction: a score-space direction
 * `direction` (nullable, same layout as scores) plus scalar parameter
 * directions (d_gap_open, d_gap_ext, d_temp).
 *   map_tangent (nullable)   directional derivative of the map:
 *                            direction only            -> HVP  (H . v)
 *                            d_gap_open = 1 only       -> dmap/dgap_open, etc.
 *   contract[B*3] (nullable) per pair, the derivatives along `direction` of
 *                            (dS/dgap_open, dS/dgap_ext, dS/dtemp); summed over
 *                            pairs these are the parameter gradients of
 *                            <cotangent, map> with cotangent = direction.
 *   value, map (nullable)    primal value and map, computed on the way.
 * Replaces: sw_affine_marginals_hvp, sw_affine_marginals_backward,
 * sw_affine_marginals_grad_{gap_open,gap_ext,temp}, soft_sw_affine_hvp,
 * soft_sw_affine_param_jacobian, soft_sw_affine_backward_full
 * (src/sw_affine/torch_hip.cpp:508-678) and the launchers sw_affine_hvp_hip /
 * sw_affine_param_grad_hip (kernels_gpu.hiph). */
ohm_status ohm_sw_affine_tangent(const float* scores, const int32_t* lengths,
                                 const int64_t* offsets, int32_t B, int32_t L1max,
                                 int32_t L2max, float gap_open, float gap_ext, float temp,
                                 const float* direction, float d_gap_open, float d_gap_ext,
                                 float d_temp, float* map_tangent, float* contract,
                                 float* value, float* map, void* workspace,
                                 size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------ */
/* Linear gap (one state)                                                    */
/* ------------------------------------------------------------------------ */

/* As ohm_sw_affine_forward with one gap score; gradients dS/dgap and
 * dS/dtemperature.  Replaces: sw_forward, sw_forward_t (forward part),
 * sw_value_grad_params, soft_sw_with_grads (src/sw/torch_hip.cpp) and the
 * launchers sw_regular_forward_hip / sw_regular_backward_hip
 * (src/sw/kernels_gpu.hip:1145-1299). */
ohm_status ohm_sw_forward(const float* scores, const int32_t* lengths, const int64_t* offsets,
                          int32_t B, int32_t L1max, int32_t L2max, float gap, float temp,
                          float* value, float* map, float* grad_gap, float* grad_temp,
                          void* workspace, size_t workspace_bytes, void* stream);

/* As ohm_sw_affine_tangent; contract is [B*2] (gap, temperature).
 * Replaces: sw_marginals_hvp, sw_marginals_backward, sw_marginals_grad_gap,
 * sw_marginals_grad_temp and the launchers sw_regular_hvp_hip /
 * sw_regular_param_grad_hip (src/sw/kernels_gpu.hip:1301-1422). */
ohm_status ohm_sw_tangent(const float* scores, const int32_t* lengths, const int64_t* offsets,
                          int32_t B, int32_t L1max, int32_t L2max, float gap, float temp,
                          const float* direction, float d_gap, float d_temp,
                          float* map_tangent, float* contract, float* value, float* map,
                          void* workspace, size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------ */
/* Needleman-Wunsch, affine gap (global alignment, three states M, I, D):  */
/* boundary states M(0,0) = 0, I(i,0) = go + (i-1) ge, D(0,j) = go + (j-1) */
/* ge, no restart, value LSE(M, I, D) at (L1, L2).                          */
/* ------------------------------------------------------------------------ */

/* Workspace bytes; op is OHM_OP_VALUE, OHM_OP_FORWARD or OHM_OP_TANGENT (0
 * otherwise). */
size_t ohm_nw_affine_workspace(int32_t op, int32_t B, int32_t L1max, int32_t L2max,
                               int64_t total_cells);

/* Value, map and scalar-parameter gradients, arguments as
 * ohm_sw_affine_forward (the gradients need the map buffer).  Pairs with an
 * empty side take the closed form: value go + (n-1) ge over the n > 0
 * characters of the other side (0 for two empty sides), zero map.
 * Replaces: nw_affine_forward, nw_affine_value_grad_params,
 * soft_nw_affine_with_grads (src/nw_affine/registry.cpp:38-46) and the
 * launchers nw_affine_forward_hip / nw_affine_backward_hip
 * (src/nw_affine/kernels_gpu.hip). */
ohm_status ohm_nw_affine_forward(const float* scores, const int32_t* lengths,
                                 const int64_t* offsets, int32_t B, int32_t L1max,
                                 int32_t L2max, float gap_open, float gap_ext, float temp,
                                 float* value, float* map, float* grad_open, float* grad_ext,
                                 float* grad_temp, void* workspace, size_t workspace_bytes,
                                 void* stream);

/* Forward-over-reverse pass of the Needleman-Wunsch affine model, arguments
 * and outputs as ohm_sw_affine_tangent (op OHM_OP_TANGENT for the workspace
 * query): the HVP (direction), the map's parameter sensitivities (unit
 * d_gap_open / d_gap_ext / d_temp) and, with contract [B*3], the map VJP's
 * parameter parts.  Replaces: nw_affine_marginals_hvp,
 * nw_affine_marginals_grad_*, nw_affine_marginals_backward,
 * soft_nw_affine_{hvp,param_jacobian,backward_full}
 * (src/nw_affine/registry.cpp:33-46) and nw_affine_hvp_cpu /
 * nw_affine_param_grad_cpu (src/nw_affine/kernels_cpu.cpp:343-1002). */
ohm_status ohm_nw_affine_tangent(const float* scores, const int32_t* lengths,
                                 const int64_t* offsets, int32_t B, int32_t L1max,
                                 int32_t L2max, float gap_open, float gap_ext, float temp,
                                 const float* direction, float d_gap_open, float d_gap_ext,
                                 float d_temp, float* map_tangent, float* contract, float* value,
                                 float* map, void* workspace, size_t workspace_bytes,
                                 void* stream);

/* ------------------------------------------------------------------------ */
/* Saigo-Vert, affine gap (local alignment counting every monotone matched- */
/* pair skeleton once): D also opens from I (no D -> I edge), paths end in */
/* M only, S = LSE(0, {M(i,j)}).                                           */
/* ------------------------------------------------------------------------ */

/* Workspace bytes; op is OHM_OP_VALUE, OHM_OP_FORWARD or OHM_OP_TANGENT (0
 * otherwise). */
size_t ohm_sv_affine_workspace(int32_t op, int32_t B, int32_t L1max, int32_t L2max,
                               int64_t total_cells);

/* Value, map and scalar-parameter gradients, arguments as
 * ohm_sw_affine_forward (the gradients need the map buffer); a pair with an
 * empty side has value 0 (the empty alignment).  Replaces: sv_affine_forward,
 * sv_affine_value_grad_params, soft_sv_affine_with_grads
 * (src/sv_affine/registry.cpp) and the launchers sv_affine_forward_hip /
 * sv_affine_backward_hip (src/sv_affine/kernels_gpu.hip). */
ohm_status ohm_sv_affine_forward(const float* scores, const int32_t* lengths,
                                 const int64_t* offsets, int32_t B, int32_t L1max,
                                 int32_t L2max, float gap_open, float gap_ext, float temp,
                                 float* value, float* map, float* grad_open, float* grad_ext,
                                 float* grad_temp, void* workspace, size_t workspace_bytes,
                                 void* stream);

/* ------------------------------------------------------------------------ */
/* Linear-gap Needleman-Wunsch and Saigo-Vert: the affine models with one   */
/* gap score in both roles (each move sequence has exactly one state        */
/* sequence, so they are the same function of the inputs).                 */
/* ------------------------------------------------------------------------ */

size_t ohm_nw_workspace(int32_t op, int32_t B, int32_t L1max, int32_t L2max,
                        int64_t total_cells);
size_t ohm_sv_workspace(int32_t op, int32_t B, int32_t L1max, int32_t L2max,
                        int64_t total_cells);

/* Value, map, dS/dgap and dS/dtemperature, arguments as ohm_sw_forward.
 * Replaces: nw_forward, nw_value_grad_params, soft_nw_with_grads
 * (src/nw/registry.cpp:35-49) and nw_forward_cpu / nw_backward_cpu
 * (src/nw/kernels_cpu.cpp:78-272). */
ohm_status ohm_nw_forward(const float* scores, const int32_t* lengths, const int64_t* offsets,
                          int32_t B, int32_t L1max, int32_t L2max, float gap, float temp,
                          float* value, float* map, float* grad_gap, float* grad_temp,
                          void* workspace, size_t workspace_bytes, void* stream);

/* As ohm_nw_forward for Saigo-Vert.  Replaces: sv_linear_forward,
 * sv_linear_value_grad_params, soft_sv_linear_with_grads
 * (src/sv_linear/registry.cpp) and sv_linear_forward_cpu /
 * sv_linear_backward_cpu (src/sv_linear/kernels_cpu.cpp:94-365). */
ohm_status ohm_sv_forward(const float* scores, const int32_t* lengths, const int64_t* offsets,
                          int32_t B, int32_t L1max, int32_t L2max, float gap, float temp,
                          float* value, float* map, float* grad_gap, float* grad_temp,
                          void* workspace, size_t workspace_bytes, void* stream);

/* Forward-over-reverse pass of the Saigo-Vert affine model, as
 * ohm_nw_affine_tangent.  Replaces: sv_affine_marginals_hvp,
 * sv_affine_marginals_grad_*, sv_affine_marginals_backward,
 * soft_sv_affine_{hvp,param_jacobian,backward_full} and sv_affine_hvp_cpu /
 * sv_affine_param_grad_cpu (src/sv_affine/kernels_cpu.cpp:362-935). */
ohm_status ohm_sv_affine_tangent(const float* scores, const int32_t* lengths,
                                 const int64_t* offsets, int32_t B, int32_t L1max,
                                 int32_t L2max, float gap_open, float gap_ext, float temp,
                                 const float* direction, float d_gap_open, float d_gap_ext,
                                 float d_temp, float* map_tangent, float* contract, float* value,
                                 float* map, void* workspace, size_t workspace_bytes,
                                 void* stream);

/* ------------------------------------------------------------------------ */
/* Host (CPU) twins: same arguments on host memory, no workspace or stream. */
/* `threads` <= 0 uses every hardware thread.  Replace the reference CPU     */
/* dispatch (src/sw_affine/torch_cpu.cpp:593-707, src/sw/torch_cpu.cpp:522-591) */
/* and kernels (src/sw/kernels_cpu.cpp, src/sw_affine/kernels_cpu.cpp).       */
/* ------------------------------------------------------------------------ */

ohm_status ohm_sw_affine_forward_cpu(const float* scores, const int32_t* lengths,
                                     const int64_t* offsets, int32_t B, int32_t L1max,
                                     int32_t L2max, float gap_open, float gap_ext, float temp,
                                     float* value, float* map, float* grad_open,
                                     float* grad_ext, float* grad_temp, int32_t threads);

ohm_status ohm_sw_affine_tangent_cpu(const float* scores, const int32_t* lengths,
                                     const int64_t* offsets, int32_t B, int32_t L1max,
                                     int32_t L2max, float gap_open, float gap_ext, float temp,
                                     const float* direction, float d_gap_open,
                                     float d_gap_ext, float d_temp, float* map_tangent,
                                     float* contract, float* value, float* map,
                                     int32_t threads);

/* Replaces nw_affine_forward_cpu / nw_affine_backward_cpu
 * (src/nw_affine/kernels_cpu.cpp:100-335). */
ohm_status ohm_nw_affine_forward_cpu(const float* scores, const int32_t* lengths,
                                     const int64_t* offsets, int32_t B, int32_t L1max,
                                     int32_t L2max, float gap_open, float gap_ext, float temp,
                                     float* value, float* map, float* grad_open,
                                     float* grad_ext, float* grad_temp, int32_t threads);

/* Replaces sv_affine_forward_cpu / sv_affine_backward_cpu
 * (src/sv_affine/kernels_cpu.cpp:96-356). */
ohm_status ohm_sv_affine_forward_cpu(const float* scores, const int32_t* lengths,
                                     const int64_t* offsets, int32_t B, int32_t L1max,
                                     int32_t L2max, float gap_open, float gap_ext, float temp,
                                     float* value, float* map, float* grad_open,
                                     float* grad_ext, float* grad_temp, int32_t threads);

ohm_status ohm_nw_forward_cpu(const float* scores, const int32_t* lengths,
                              const int64_t* offsets, int32_t B, int32_t L1max, int32_t L2max,
                              float gap, float temp, float* value, float* map, float* grad_gap,
                              float* grad_temp, int32_t threads);

ohm_status ohm_sv_forward_cpu(const float* scores, const int32_t* lengths,
                              const int64_t* offsets, int32_t B, int32_t L1max, int32_t L2max,
                              float gap, float temp, float* value, float* map, float* grad_gap,
                              float* grad_temp, int32_t threads);

ohm_status ohm_nw_affine_tangent_cpu(const float* scores, const int32_t* lengths,
                                     const int64_t* offsets, int32_t B, int32_t L1max,
                                     int32_t L2max, float gap_open, float gap_ext, float temp,
                                     const float* direction, float d_gap_open, float d_gap_ext,
                                     float d_temp, float* map_tangent, float* contract,
                                     float* value, float* map, int32_t threads);

ohm_status ohm_sv_affine_tangent_cpu(const float* scores, const int32_t* lengths,
                                     const int64_t* offsets, int32_t B, int32_t L1max,
                                     int32_t L2max, float gap_open, float gap_ext, float temp,
                                     const float* direction, float d_gap_open, float d_gap_ext,
                                     float d_temp, float* map_tangent, float* contract,
                                     float* value, float* map, int32_t threads);

ohm_status ohm_sw_forward_cpu(const float* scores, const int32_t* lengths,
                              const int64_t* offsets, int32_t B, int32_t L1max, int32_t L2max,
                              float gap, float temp, float* value, float* map, float* grad_gap,
                              float* grad_temp, int32_t threads);

ohm_status ohm_sw_tangent_cpu(const float* scores, const int32_t* lengths,
                              const int64_t* offsets, int32_t B, int32_t L1max, int32_t L2max,
                              float gap, float temp, const float* direction, float d_gap,
                              float d_temp, float* map_tangent, float* contract, float* value,
                              float* map, int32_t threads);

/* ------------------------------------------------------------------------ */
/* Input scan (drop-in validation)                                           */
/* ------------------------------------------------------------------------ */

/* One pass over padded scores [B, L1, L2] (every cell of every box) and the
 * lengths [B, 2] (nullable) for the checks the reference's Python engine and
 * bindings make before a call (orihime/operator.py:249-386,
 * src/sw_affine/torch_hip.cpp:69-99), which it runs as several torch
 * reductions with host copies.  Writes counts[4] (device int32): [0] +inf
 * cells, [1] NaN cells and finite cells with |x| > bound other than
 * `marker` (the legacy finite mask value), [2] -inf cells, [3] lengths
 * entries outside [0, L1] x [0, L2]; and pair_max[B] (device float): each
 * box's largest non-NaN cell, -inf for an empty box (the mask sentinel's
 * reference point).  Enqueues three small launches on `stream`; nothing
 * needs clearing first and nothing synchronises the host. */
ohm_status ohm_scan_inputs(const float* scores, const int32_t* lengths, int32_t B, int32_t L1,
                           int32_t L2, float bound, float marker, int32_t* counts,
                           float* pair_max, void* stream);

/* ------------------------------------------------------------------------ */
/* Allocation-free entry point                                               */
/* ------------------------------------------------------------------------ */

/* Bytes of the library's fixed device workspace (64 MiB per GPU, held from
 * the first allocation-free call on that GPU until the process ends). */
size_t ohm_scratch_bytes(void);

/* ohm_sw_affine_forward without a workspace argument: the sweeps use the
 * library's fixed device array instead, so the call allocates nothing and
 * takes no caller buffer beyond inputs and outputs.  No counterpart in the
 * reference; it is for callers that cannot allocate (code under HIP graph
 * capture, serving loops or harnesses that forbid hipMalloc).
 * total_cells: sum_b L1_b * L2_b for the packed layout (offsets != NULL; an
 * upper bound is accepted), ignored for the padded layout.  A batch whose
 * workspace (ohm_sw_affine_workspace) exceeds ohm_scratch_bytes() is
 * rejected with OHM_EWORKSPACE.  All calls on one GPU share the array:
 * calls that overlap on one GPU (two streams, two threads) race.  The first
 * call on each GPU looks the array up (hipGetSymbolAddress on the current
 * device); make it outside graph capture. */
ohm_status ohm_sw_affine_forward_noalloc(const float* scores, const int32_t* lengths,
                                         const int64_t* offsets, int64_t total_cells, int32_t B,
                                         int32_t L1max, int32_t L2max, float gap_open,
                                         float gap_ext, float temp, float* value, float* map,
                                         float* grad_open, float* grad_ext, float* grad_temp,
                                         void* stream);

#ifdef __cplusplus
}
#endif

#endif /* OHM_SW_H */
