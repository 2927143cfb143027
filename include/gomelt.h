/*
 * gomelt.h -- C ABI of libgomelt.so, the MI355X (gfx950) implementation of
 * GO-MELT's multilevel explicit finite-element heat step.
 *
 * The reference has no FFI: its boundary is the Python module surface of
 * go_melt/computeFunctions.py, imported by go_melt/go_melt.py:10.  Each entry
 * point below replaces one function (or one fused group of functions) of that
 * surface; the comment above it cites the reference lines it stands in for.
 * The Python facade go-melt_amd/computeFunctions.py binds these with ctypes
 * and keeps the reference's names, argument order and return tuples.
 *
 * Conventions
 *  - Every function returns 0 on success and a negative code on failure;
 *    gm_last_error() returns the message of the last failure on this thread.
 *  - All array arguments are device pointers (allocated with gm_malloc or any
 *    HIP allocator) unless the name ends in _host.  Fields are FP32, node
 *    numbering is x-fastest: i = ix + nx*(iy + ny*iz)  (reference cF:645-689).
 *  - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *    Calls are asynchronous with respect to the host unless stated.
 *  - The library never frees caller buffers.
 */
#ifndef GOMELT_H
#define GOMELT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Plain data                                                                */
/* ------------------------------------------------------------------------ */

/* A uniform structured hexahedral mesh: n nodes per axis, node i of an axis
 * at x0 + i*h (reference createMesh3D, cF:32-64; spacing from calc_length_h,
 * cF:1916-1942). */
typedef struct gm_grid {
    int32_t n[3];
    float x0[3];
    float h[3];
} gm_grid;

/* Material, laser and surface constants, already in the units the kernels
 * use (reference SetupProperties, cF:267-345).  k values in W/(m K) (divided
 * by 1000 in the kernels, cF:2602), cp in J/(kg K), rho in kg/mm^3,
 * h_conv in W/(m^2 K) (the input value times 1e6, cF:319).  When k_table_n
 * (cp_table_n) is > 0 a piecewise-linear table of that many knots replaces
 * the linear bulk-k (solid-cp) law; this is an extension, the reference model
 * is the case k_table_n = cp_table_n = 0. */
#define GM_MAX_KNOTS 16
typedef struct gm_material {
    float T_amb, T_solidus, T_liquidus, T_boiling;
    float k_powder, k_bulk_a0, k_bulk_a1, k_fluid;
    float cp_solid_a0, cp_solid_a1, cp_mushy, cp_fluid;
    float rho;
    int32_t k_table_n, cp_table_n;
    float k_table_T[GM_MAX_KNOTS], k_table_v[GM_MAX_KNOTS];
    float cp_table_T[GM_MAX_KNOTS], cp_table_v[GM_MAX_KNOTS];
    float h_conv, sigma_eps;          /* sigma_sb * emissivity               */
    float evc_cp, CT, CM, Lev;        /* evaporation model (cF:337-343)      */
    float laser_eta, laser_radius, laser_depth;
} gm_material;

/* Static description of a three-level run (reference SetupLevels cF:115-264,
 * SetupNonmesh cF:348-419, getStaticSubcycle cF:473-492). */
typedef struct gm_config {
    gm_grid level[4];          /* [1..3] initial L1/L2/L3 meshes; [0] Level0 */
    float l1_bc[5];            /* L1 Dirichlet values: y0, y1, x0, x1, z0     */
    float layer_height;        /* mm                                          */
    int32_t n2, n3;            /* subcycle_num_L2, subcycle_num_L3            */
    int32_t ratio12[3];        /* L1L2Eratio (go_melt.py:42-44)               */
    int32_t ratio23[3];        /* L2L3Eratio (go_melt.py:46-48)               */
    float move_lo[2][3];       /* window move limits of L2 and L3 (cF:1886)   */
    float move_hi[2][3];
    int32_t ov_lo[4][3];       /* original overlap node boxes: [2] L1 nodes in
                                  L2, [3] L2 nodes in L3, [0] Level0 nodes in
                                  L3, [1] Level0 nodes in L2 (strided)        */
    int32_t ov_n[4][3];
    float ov_x0[4][3];         /* original overlap coordinates (first node)   */
    int32_t l0_stride_l2[3];   /* Level0 node stride of the L2 overlap        */
    int32_t l0_layer_delta;    /* Level0 layer_idx_delta                      */
} gm_config;

/* ------------------------------------------------------------------------ */
/* Device, memory and streams                                                */
/* ------------------------------------------------------------------------ */

const char* gm_last_error(void);
int gm_device_init(int device);
int gm_device_name(char* buf, int len);
int gm_malloc(void** ptr, size_t bytes);
int gm_free(void* ptr);
int gm_memcpy_h2d(void* dst, const void* src_host, size_t bytes, void* stream);
int gm_memcpy_d2h(void* dst_host, const void* src, size_t bytes, void* stream);
int gm_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream);
int gm_memset(void* dst, int value, size_t bytes, void* stream);
int gm_stream_create(void** stream);
int gm_stream_destroy(void* stream);
int gm_stream_sync(void* stream);
int gm_event_create(void** ev);
int gm_event_destroy(void* ev);
int gm_event_record(void* ev, void* stream);
int gm_stream_wait_event(void* stream, void* ev); /* later work on stream waits for ev */
int gm_event_elapsed_ms(void* start, void* stop, float* ms);

/* ------------------------------------------------------------------------ */
/* Single operators (one reference function each)                            */
/* ------------------------------------------------------------------------ */

/* computeStateProperties (cF:2567-2614): phase flags and nodal k (W/mm K),
 * rho*cp.  S1 in/out is FP32 0/1 (S1_out may alias S1); S2_out is 0/1 bytes. */
int gm_state_properties(const gm_material* mat, const float* T, const float* S1,
                        int64_t nn, int64_t n_substrate, float* S1_out,
                        uint8_t* S2_out, float* k_out, float* rhocp_out, void* stream);

/* solveMatrixFreeFE (cF:582-642): T+ = [sum_e (diag M_e - K_e) T_e + F + Corr]
 * / sum_e M_e over elements 0..ne_active-1 with element-mean k and rho*cp.
 * F and Corr may be NULL (zero).  Nodes in no active element divide 0/0. */
int gm_solve_matrix_free(const gm_grid* grid, int64_t ne_active, const float* k,
                         const float* rhocp, float dt, const float* T, const float* F,
                         const float* Corr, float* T_out, void* stream);

/* computeSourcesL3 (cF:2960-3012): nodal Gaussian source of one laser
 * position v = (x, y, z) and power P, integrated with 2x2x2 Gauss points. */
int gm_source_l3(const gm_grid* grid, const gm_material* mat, const float* v_host,
                 float P, float* F_out, void* stream);

/* The same source averaged over n_rows laser rows (row stride 7 floats: x, y,
 * z, Ljump, Ldwell, dt, P), F = sum_r (1/n_rows) F_r in row order; the mean
 * source the L3 solves of an L2 substep see (cF:3367-3412), for the energy
 * audit of a jump over them.  Not a reference function. */
int gm_source_l3_mean(const gm_grid* grid, const gm_material* mat, const float* rows_host,
                      int n_rows, float* F_out, void* stream);

/* computeLevelSource (cF:2667-2730): coarse source averaged over n_pos laser
 * rows (row stride 7 floats: x, y, z, Ljump, Ldwell, dt, P), evaluated at the
 * fine mesh's Gauss points and restricted with the coarse shape functions. */
int gm_level_source(const gm_grid* coarse, const gm_grid* fine, const gm_material* mat,
                    const float* rows_host, int n_pos, float* F_out, void* stream);

/* Melt-pool size of a level's temperature field.  No reference counterpart:
 * the reference only forms S2 = T >= T_liquidus (cF:2585); this is the
 * measure the surrogate's error is taken in.  dims: 8 floats of device memory,
 * written asynchronously on `stream`:
 *   [0] depth  = max over vertical edges with a molten upper node of the
 *                distance from the top node plane down to the T_liquidus
 *                crossing (linear along the edge; a molten bottom node counts
 *                the full mesh depth),
 *   [1] width  = y extent of the molten part of the top node plane, its ends
 *                at the T_liquidus crossings of y edges (or on the mesh
 *                boundary where a boundary node is molten),
 *   [2] length = the same along x,
 *   [3] number of nodes with T >= T_liquidus,
 *   [4..7] y lo, y hi, x lo, x hi of the molten top region (NaN if none).
 * Sizes are 0 where nothing is molten. */
int gm_melt_pool(const gm_grid* grid, const float* T, float T_liquidus, float* dims,
                 void* stream);

/* Energy balance of one explicit update T -> T_new of a level, the audit
 * that decides whether an update other than the plain step (a surrogate) is
 * kept.  No reference counterpart; it checks the discrete identity of
 * solveMatrixFreeFE (cF:582-642): with the lumped capacity
 * C_i = sum_e rhocp_e V_e / 8 and the element-mean-k stiffness K (K times a
 * constant is 0), the unconstrained update Tt = T + dt (F + Corr - K T) / C
 * satisfies sum_i C_i (Tt - T)_i = dt sum_i (F + Corr)_i.  k, rhocp, F, Corr
 * have the meaning and units of gm_solve_matrix_free (F, Corr nullable).
 * Nodes on the Dirichlet faces in bc_faces (bits 0..4: x = 0, x = max,
 * y = 0, y = max, z = 0) and nodes held at the floor (T_new <= T_clamp while
 * Tt < T_clamp; pass -INFINITY for no floor) count with Tt; all others with
 * T_new.  Nodes in no active element are skipped.  sums: 8 doubles of device
 * memory, written asynchronously on `stream`:
 *   [0] sum over free nodes of C (T_new - T)      [1] held nodes, C (Tt - T)
 *   [2] dt sum (F + Corr)                          [3] dt sum K T  (~ 0)
 *   [4] held nodes, C (T_new - Tt) (energy the constraints add)
 *   [5] scale: sum |dt (F + Corr)| + sum |C dT|    [6] residual ([0]+[1]-[2])/[5]
 *   [7] number of held nodes.
 * Per-node terms and sums in FP64. */
int gm_energy_balance(const gm_grid* grid, int64_t ne_active, const float* k,
                      const float* rhocp, float dt, const float* T, const float* F,
                      const float* Corr, const float* T_new, int bc_faces, float T_clamp,
                      double* sums, void* stream);

/* computeConvRadBC (cF:2207-2301): adds the top-face convection, radiation
 * and evaporation flux of the top layer of elements 0..ne_active-1 to F. */
int gm_conv_rad_bc(const gm_grid* grid, const gm_material* mat, const float* T,
                   int64_t ne_active, float* F, void* stream);

/* interpolatePoints (cF:1131-1210): trilinear interpolation of u (on src) at
 * the nodes of the tensor grid `target`; a point whose weights leave
 * [-0.01, 1.01] gets 0. */
int gm_interpolate_points(const gm_grid* src, const float* u, const gm_grid* target,
                          float* out, void* stream);

/* interpolatePointsMatrix (cF:1028-1107): the weights (n, 8) and coarse node
 * ids (n, 8) that gm_interpolate_points applies. */
int gm_interpolate_matrix(const gm_grid* src, const gm_grid* target, float* Nc,
                          int32_t* node, void* stream);

/* interpolate_w_matrix (cF:1110-1128). */
int gm_interpolate_w_matrix(const float* Nc, const int32_t* node, int64_t n,
                            const float* T, float* out, void* stream);

/* Restriction of fine-level T' terms onto a coarse level (the BCOO products
 * of cF:1477-1565 / 2733-2810 / 2857-2914 for flux, cF:1396-1474 / 3057-3132 /
 * 3169-3221 for mass):
 *   V_A += -sum_e sum_q [grad N_A . k_e grad Tp + N_A rhocp_e (TpA - TpB)/dt] w
 * k_e, rhocp_e are element means of the nodal properties of (T, S1) with
 * n_substrate forced solid.  Tp (flux) or TpA/TpB (mass) may be NULL to drop
 * that term.  Sums are accumulated in FP64.  V has the coarse node count. */
int gm_restrict(const gm_grid* coarse, const gm_grid* fine, const gm_material* mat,
                const float* T, const float* S1, int64_t n_substrate, const float* Tp,
                const float* TpA, const float* TpB, float dt, float* V, void* stream);

/* gm_restrict without device allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as gm_restrict, but the descriptor, axis
 * tables, FP64 element partials and node box live in a static device array of
 * 16 MiB per GPU (held until the process ends), so the call allocates and
 * frees nothing, copies nothing from the host and does not synchronise: the
 * result is ready when `stream` is.  A pair whose scratch would exceed 16 MiB
 * is rejected with an error before any launch (use gm_restrict for it).  All
 * calls on one GPU share that GPU's array: calls that overlap in time on one
 * GPU (other streams, other threads) race on it.  Like gm_restrict, V is
 * added to, so the caller clears it first. */
int gm_restrict_noalloc(const gm_grid* coarse, const gm_grid* fine, const gm_material* mat,
                        const float* T, const float* S1, int64_t n_substrate, const float* Tp,
                        const float* TpA, const float* TpB, float dt, float* V, void* stream);

/* getNewTprime (cF:2060-2099): CoarseT[overlap box] <- fine field interpolated
 * at those coarse nodes; then Tprime = FineT - prolong(CoarseT). */
int gm_new_tprime(const gm_grid* fine, const float* FineT, const gm_grid* coarse,
                  float* CoarseT, const int32_t* ov_lo_host, const int32_t* ov_n_host,
                  float* Tprime, void* stream);

/* ------------------------------------------------------------------------ */
/* The time loop: a device-resident three-level state                        */
/* ------------------------------------------------------------------------ */

typedef struct gm_engine gm_engine;

/* Fields exposed by gm_engine_field. */
enum {
    GM_T0 = 0,        /* FP32 temperature (levels 1-3)                       */
    GM_TPRIME0 = 1,   /* FP32 T' (levels 2-3)                                */
    GM_S1 = 2,        /* level 1: FP32; levels 2-3: uint8; Level0: uint8     */
    GM_S2 = 3,        /* level 3: uint8                                      */
    GM_ACCUM = 4,     /* Level0 FP32 accumulated melt time                   */
    GM_MAX_ACCUM = 5, /* Level0 FP32 max accumulated melt time               */
    GM_HIST_L2 = 6,   /* corrector L2 T after L2 substep `level` (L2all)     */
    GM_HIST_L3 = 7,   /* corrector L3 T after L2 substep `level` (L3all)     */
    GM_HIST_TP3 = 8,  /* corrector L3 T' after L2 substep `level` (L3pall)   */
    GM_ACCUM_L3 = 9,  /* FP32 melt time on the L3 window nodes (level ignored)*/
    GM_MAX_ACCUM_L3 = 10, /* FP32 max melt time on the L3 window nodes        */
    GM_RESET_L3 = 11  /* uint8 reset mask of the last gm_engine_step         */
};

/* Buffers of a staged subcycle between its l2 and l3 stages (all FP32 with the
 * L3 node count, except GM_STAGE_S1_IN, uint8). */
enum {
    GM_STAGE_L3_IN = 0,    /* L3 field at the start of the L2 substep           */
    GM_STAGE_L3_OUT = 1,   /* where the substep's final L3 field goes           */
    GM_STAGE_FACE_NEW = 2, /* L3 face values from the substep's new L2 field    */
    GM_STAGE_FACE_OLD = 3, /* ... from its start-of-substep L2 field            */
    GM_STAGE_S1_IN = 4,    /* L3 phase flags at the start of the substep        */
    GM_STAGE_L3_LAST = 5   /* the final L3 field of the last finished substep   */
};

/* SetupLevels (cF:115-264) + getStaticNodesAndElements/getStaticSubcycle:
 * allocates all level state on the current device and initialises it to
 * T_amb, S1 = 0.  `stream` is kept and used by every engine call. */
int gm_engine_create(const gm_config* cfg, const gm_material* mat, void* stream,
                     gm_engine** out);
int gm_engine_destroy(gm_engine* e);

/* Device pointer and element count of a field (see the enum above). */
int gm_engine_field(gm_engine* e, int level, int field, void** ptr, int64_t* count);

/* Current mesh of a level (1..3) after the last window move. */
int gm_engine_grid(gm_engine* e, int level, gm_grid* out);

/* Static per-layer quantities: calcStaticTmpNodesAndElements (cF:495-517)
 * and getSubstrateNodes (cF:562-579), given by the caller. */
int gm_engine_set_layer(gm_engine* e, int64_t l1_ne_active, int64_t l1_nn_active,
                        const int64_t* substrate4_host);

/* moveEverything (cF:2400-2510) for laser row v and start position vstart
 * (3 floats each): moves L3 and L2 with the laser, re-interpolates T0/T' onto
 * the moved meshes, pulls S1/S2 from Level0.  move_v_out (3 ints, host)
 * receives the L2 move in L2 elements. */
int gm_engine_move(gm_engine* e, const float* v_host, const float* vstart_host,
                   int32_t* move_v_out_host);

/* subcycleGOMELT (cF:3224-3632) for N2*N3 toolpath rows (row stride 7):
 * one L1 step with L2/L3 subcycling, predictor and corrector, melt-time
 * accumulation on Level0 and the Level0 S1/S2 update.  The first call with a
 * given layer state records a hipGraph; later calls replay it. */
int gm_engine_subcycle(gm_engine* e, const float* rows_host, int n_rows);

/* subcycleGOMELT (cF:3224-3632) stage by stage, for a caller that advances the
 * L3 window of some L2 substeps itself (the learned operator in place of the
 * N3 L3 solves of cF:3367-3412 and cF:3530-3590).  stage_begin takes the rows
 * and runs the prologue; then, 2*N2 times (Part 1, then Part 2), stage_l2 runs
 * the L2 substep (restriction, L2 solve, L3 face values; part, k returned) and
 * stage_l3 finishes it: external = 0 runs the N3 L3 solves, external = 1
 * takes the final L3 field the caller wrote into GM_STAGE_L3_OUT (phase
 * flags, melt time over the substep, carried k and cp), then T'.  The
 * GM_STAGE_L3_OUT buffer is never the start field's, so the plain solves can
 * still run after a caller wrote it; their final field (and, after an
 * external stage, the caller's) is GM_STAGE_L3_LAST once the stage is done.  The l3
 * stage closing Part 1 runs the L1 corrector, the one closing Part 2 the
 * epilogue.  With external = 0 everywhere the result equals
 * gm_engine_subcycle's; kernels are launched directly, not from a graph. */
int gm_engine_stage_begin(gm_engine* e, const float* rows_host, int n_rows);
int gm_engine_stage_l2(gm_engine* e, int32_t* part_host, int32_t* k_host);
int gm_engine_stage_l3(gm_engine* e, int external);
int gm_engine_stage_field(gm_engine* e, int which, void** ptr, int64_t* count);

/* Many independent three-level runs advanced together (the batched
 * configuration: each instance its own L1/L2/L3 stack, subcycleGOMELT
 * cF:3224-3632 per instance).  The engines run their subcycles stage by
 * stage in lockstep, each on its own stream (several may share one); the
 * batched solves run on the first engine's stream, after the other streams'
 * L2 stages and before their L3 closes (events).  Stages: (gm_engine_stage_begin and
 * gm_engine_stage_l2 on each); gm_multi_l3 then runs the N3 L3 solves of the
 * current L2 substep of every engine as N3 launches of the batched step (all
 * engines' bricks in one launch per solve, melt time in Part 2) and finishes
 * each engine's L3 stage as gm_engine_stage_l3(e, 0) would.  Every engine
 * must be between the l2 and l3 stages of the same (part, k), with equal N2,
 * N3 and the same kind of property tables.  No reference counterpart (the
 * reference runs one instance per process). */
typedef struct gm_multi gm_multi;
int gm_multi_create(gm_engine* const* engines, int n, gm_multi** out);
int gm_multi_l3(gm_multi* m);
/* The whole subcycle of every engine after gm_engine_stage_begin on each:
 * for each of the 2*N2 L2 substeps the L2 stage of every engine, then
 * gm_multi_l3.  Gives what the caller's loop over gm_engine_stage_l2 and
 * gm_multi_l3 gives (to FP32 rounding; see gm_multi_set_batch_l2), without a
 * host call per engine and substep.  It does not wait for the GPU: the host
 * may run up to 32 substeps ahead (pinned item slots), and each engine's
 * stream is ordered after the batched work when the call returns. */
int gm_multi_subcycle(gm_multi* m);
/* Brick shape (rows per band x planes, as gm_batch_set_tile) of the batched
 * L3 solves of both parts, e.g. the one gm_batch_autotune picks for the same
 * L3 windows (default 1 x 3).  Synchronises the first engine's stream. */
int gm_multi_set_tile(gm_multi* m, int yb, int zc);
/* With on != 0 (the default) gm_multi_subcycle also batches, per substep,
 * the restrictions of the L3 T' terms, the L2 solves, the L3 face values and
 * T' + injection of all engines (one launch each on the first engine's
 * stream; an engine whose restriction is not of the node-aligned form with
 * carried properties restricts on its own stream), when every engine is on
 * the node-aligned path; otherwise, or with on = 0, those run per engine as
 * in gm_engine_stage_l2 / _l3. */
int gm_multi_set_batch_l2(gm_multi* m, int on);
/* With on = 0 the batched substep leaves each engine's restriction of the
 * L3 T' terms on its own stream (the streams joined around it every
 * substep), as for an engine whose restriction the batched launch does not
 * run; default on. */
int gm_multi_set_batch_restrict(gm_multi* m, int on);
int gm_multi_destroy(gm_multi* m);

/* stepGOMELT (cF:2304-2397) for one toolpath row.  Writes the reference's
 * return value _resetmask (cF:2394-2397: L3 nodes whose S2 went 0 -> 1) into
 * the GM_RESET_L3 field.  In the engine's own melt-time mode (the default,
 * see gm_engine_set_accum_mode) it also applies the single-step driver's
 * melt-time bookkeeping (go_melt.py:339-357) to the Level0 maps. */
int gm_engine_step(gm_engine* e, const float* row_host);

/* Engine state for the driver's per-layer checkpoint (go_melt.py:390-411)
 * and its resume (go_melt.py:115-129): every persistent array of the three
 * levels (T0, T', S1, S2, the L1 S1 storage rows, the L3 melt-time window)
 * and the window positions and overlap boxes, in one host blob of
 * gm_engine_state_bytes bytes.  Level0 (S1 and the melt-time maps) is not in
 * it: the checkpoint stores those as the reference does (Levels[0],
 * accum_time, max_accum_time).  load_state accepts only a blob saved by an
 * engine of the same deck (gm_config).  Both are synchronous. */
int gm_engine_state_bytes(gm_engine* e, int64_t* bytes);
int gm_engine_save_state(gm_engine* e, void* host);
int gm_engine_load_state(gm_engine* e, const void* host, int64_t bytes);

/* Who keeps the melt-time maps.  external = 0 (default): the engine keeps
 * them on its Level0 (GM_ACCUM, GM_MAX_ACCUM); gm_engine_subcycle gathers
 * the L3 window from them and scatters it back, gm_engine_step updates them.
 * external = 1: the caller keeps them, as the reference driver does
 * (go_melt.py:339-357, 436-456): before gm_engine_subcycle it writes the
 * window values max_accum_time[idx], accum_time[idx] into GM_MAX_ACCUM_L3 and
 * GM_ACCUM_L3, after it reads the updated window back from them
 * (subcycleGOMELT's inputs and returns, cF:3224-3238, 3568-3578);
 * gm_engine_step leaves melt time alone. */
int gm_engine_set_accum_mode(gm_engine* e, int external);

/* Current overlap index boxes after the last window move (moveEverything
 * cF:2477-2497 and update_overlap_nodes_coords*, cF:1727-1845): which = 0:
 * Level0 nodes of the L3 window (Levels[0]["overlapNodes"], whose flat
 * indices are Levels[0]["idx"]); 1: Level0 nodes of the L2 window, strided
 * ("overlapNodes_L2", "idx_L2"); 2: L1 nodes inside L2; 3: L2 nodes inside
 * L3.  Axis d covers lo[d] + stride[d] * (0 .. n[d]-1).  3 ints each, host. */
int gm_engine_overlap_box(gm_engine* e, int which, int32_t* lo_host, int32_t* n_host,
                          int32_t* stride_host);

/* stepGOMELTDwellTime (cF:2617-2664): L1-only step of length dt. */
int gm_engine_step_dwell(gm_engine* e, float dt);

/* Layer change of the single-step driver path (go_melt.py:198-290): shifts
 * the L1 z axis by n_layers layer heights (re-interpolating T0), swaps the S1
 * storage row, shifts Level0 down one layer and resets melt-time maxima. */
int gm_engine_layer_change(gm_engine* e, int kidx, float new_l1_z0, float laser_z,
                           int shift_level0);

/* Zero Tprime0 of L2 and L3 (dwell hand-over, go_melt.py:360-368). */
int gm_engine_clear_tprime(gm_engine* e);

/* Set Level0 S1 = 1 on its first n nodes (go_melt.py:312-313). */
int gm_engine_level0_substrate(gm_engine* e, int64_t n);

/* min/max of T0 per level and a finite/range flag (printLevelMaxMin,
 * cF:3635-3665).  Synchronous.  minmax_host: 6 floats (L1 min,max, L2, L3). */
int gm_engine_minmax(gm_engine* e, float* minmax_host);

/* Measurement helper: average duration (ms) of `iters` back-to-back launches
 * of one kernel of the L1 step on the engine's current state, recorded as a
 * hipGraph and timed with HIP events on the engine stream.  which = 3: L3
 * level step, 2: L2 level step, 1: L1 level step, 4: L3->L2 restriction
 * (these overwrite scratch buffers only); which = 0: replays of the recorded
 * subcycle graph (advances the engine state).  Synchronous. */
int gm_engine_time_kernel(gm_engine* e, int which, int iters, float* ms_per_launch);

/* Enable/disable hipGraph capture of gm_engine_subcycle (default enabled). */
int gm_engine_use_graph(gm_engine* e, int enable);

/* ------------------------------------------------------------------------ */
/* Batched instances                                                         */
/* ------------------------------------------------------------------------ */

/* Many independent process-parameter instances of an L3-type window, each of
 * its own mesh size and laser, advanced by ONE kernel launch per step.  No
 * reference counterpart (the reference runs one instance per process); each
 * instance's step is the L3 step of the subcycle (computeSolutions_L3,
 * cF:3015-3054: source cF:2960-3012, surface flux cF:2207-2301, Dirichlet
 * W/E/S/N/bottom faces cF:1598-1620, floor at T_amb), with properties
 * (cF:2567-2614) carried between steps.  The faces hold per-instance values
 * (initially T_amb; writable through gm_batch_field).  Not re-entrant. */
typedef struct gm_batch gm_batch;

/* n instances: grids[i] its mesh, mats[i] its material and laser constants
 * (all with property tables or all without).  T = T_amb, S1 = 0. */
int gm_batch_create(int n, const gm_grid* grids, const gm_material* mats, void* stream,
                    gm_batch** out);
int gm_batch_destroy(gm_batch* b);

/* One step of every instance: rows_host holds n toolpath rows (x, y, z,
 * Ljump, Ldwell, dt, P), row i for instance i.  batched = 1: one launch for
 * all instances; 0: one launch per instance (the same kernel body, for
 * checking and timing). */
int gm_batch_step(gm_batch* b, const float* rows_host, int batched);

/* Brick shape of the batched launch: yb rows per band x zc planes, one of
 * (1,2) (1,3) (1,4) (2,2) (2,3); default (1,3). */
int gm_batch_set_tile(gm_batch* b, int yb, int zc);

/* Autotuner: times `iters` batched steps for each of the five brick shapes
 * on this batch, keeps the fastest (yb_out, zc_out) and returns the five
 * times in ms_host (nullable, in the order above).  The state does not
 * advance.  Synchronous. */
int gm_batch_autotune(gm_batch* b, int iters, float* ms_host, int* yb_out, int* zc_out);

/* Average ms per step of `iters` steps from the current state, batched or
 * not (the state does not advance).  Synchronous. */
int gm_batch_time(gm_batch* b, int batched, int iters, float* ms_per_step);

/* Device view of instance i: field 0 = T (FP32), 1 = S1 (uint8), 2 = face
 * values (FP32).  After writing T or S1, call gm_batch_refresh. */
int gm_batch_field(gm_batch* b, int i, int field, void** ptr, int64_t* count);
int gm_batch_refresh(gm_batch* b);

/* After a caller has replaced instance i's T by its own update of a start
 * field T0 (an operator jump over several steps): S1 |= (T0 >= T_liquidus) |
 * (T >= T_liquidus), so a node that melted inside the jump stays melted, as
 * the per-step phase rule of the plain step keeps it (cF:2567-2614).  Call
 * gm_batch_refresh after it. */
int gm_batch_jump_state(gm_batch* b, int i, const float* T0);

/* ------------------------------------------------------------------------ */
/* Learned L3 operator: convolution kernels                                  */
/* ------------------------------------------------------------------------ */

/* 3x3x3 convolutions, stride 1, zero padding 1, of the learned L3 operator.
 * No reference counterpart (the reference has no learned part).  BF16
 * operands (uint16 bit patterns) on the MFMA matrix cores, FP32 accumulation.
 * shape = {nb, D, H, W, cin, cout} of the forward convolution; cin and cout
 * multiples of 16.  Channels last: X[((b*D + z)*H + y)*W + x][cin];
 * weights Wk[tap][cout][cin], tap = (dz+1)*9 + (dy+1)*3 + (dx+1).
 *
 * The bf16 arrays are read 4 elements at a time: 8-byte aligned pointers.
 *
 * Forward: Y[v][o] = bias[o] + sum_{tap,i} X[v + s(tap)][i] Wk[tap][o][i],
 * then max(0, .) if relu; Y FP32, Ybf (nullable) its bf16 copy. */
int gm_conv3d_fwd(const int32_t* shape_host, const uint16_t* X, const uint16_t* Wk,
                  const float* bias, int relu, float* Y, uint16_t* Ybf, void* stream);

/* The same forward pass with a caller's workspace: on volumes that give the
 * chip few output tiles (coarse U-Net levels, one window) the K sum (taps x
 * input channels) is split over blocks that write partial sums into ws,
 * which a second kernel adds in a fixed order (+ bias, ReLU).
 * gm_conv3d_fwd_ws_bytes gives the workspace size the split needs (0: no
 * split for this shape); with a smaller or null ws the call runs unsplit. */
int gm_conv3d_fwd_ws_bytes(const int32_t* shape_host, int64_t* bytes_host);
int gm_conv3d_fwd_ws(const int32_t* shape_host, const uint16_t* X, const uint16_t* Wk,
                     const float* bias, int relu, float* Y, uint16_t* Ybf, void* ws,
                     int64_t ws_bytes, void* stream);

/* Gradient of the input: dX[v][i] = sum_{tap,o} dY[v - s(tap)][o] Wk[tap][o][i]
 * (Wt_work: 27*cin*cout bf16 of caller scratch for the flipped weights). */
int gm_conv3d_bwd_data(const int32_t* shape_host, const uint16_t* dY, const uint16_t* Wk,
                       uint16_t* Wt_work, float* dX, uint16_t* dXbf, void* stream);

/* Gradient of the weights, ADDED to dW (FP32, [tap][cout][cin]):
 * dW[tap][o][i] += sum_v dY[v][o] X[v + s(tap)][i]. */
int gm_conv3d_bwd_weight(const int32_t* shape_host, const uint16_t* X, const uint16_t* dY,
                         float* dW, void* stream);

/* One Adam step on FP32 master weights w (and their bf16 copy wbf, nullable):
 * m, v moments; step >= 1 (bias corrections 1 - beta^step). */
int gm_adam(float* w, uint16_t* wbf, const float* g, float* m, float* v, int64_t n, float lr,
            float beta1, float beta2, float eps, int step, void* stream);

/* bf16 (round to nearest even) copy of FP32 values. */
int gm_f32_to_bf16(const float* a, uint16_t* b, int64_t n, void* stream);

/* Elementwise and reduction kernels of the operator's network (no reference
 * counterpart).  Channels-last tensors; activations bf16, gradients FP32.
 * coarse_shape = {nb, D, H, W, C} of the coarse grid (the fine one is
 * 2D x 2H x 2W).  dir 0: forward, 1: backward. */
/* 2x2x2 average pooling: forward bf16 fine -> bf16 coarse; backward FP32
 * coarse gradient -> FP32 fine gradient (each fine voxel gets 1/8). */
int gm_nn_pool2(int dir, const int32_t* coarse_shape_host, const void* in, void* out,
                void* stream);
/* Nearest 2x upsampling: forward bf16 coarse -> bf16 fine; backward FP32 fine
 * gradient -> FP32 coarse gradient (sum of the 8). */
int gm_nn_up2(int dir, const int32_t* coarse_shape_host, const void* in, void* out,
              void* stream);
/* Y = [A | B] along channels (bf16); the backward split of an FP32 gradient. */
int gm_nn_concat(int64_t nvox, int ca, int cb, const uint16_t* A, const uint16_t* B,
                 uint16_t* Y, void* stream);
int gm_nn_split(int64_t nvox, int ca, int cb, const float* dY, float* dA, float* dB,
                void* stream);
/* dZ = (dA + dA2) where A > 0, else 0 (dA2, A nullable); FP32 and a bf16 copy
 * (dZb nullable). */
int gm_nn_relu_bwd(int64_t n, const float* dA, const float* dA2, const float* A, float* dZ,
                   uint16_t* dZb, void* stream);
/* db[c] += sum over voxels of dZ[v][c]. */
int gm_nn_bias_grad(int64_t nvox, int C, const float* dZ, float* db, void* stream);
/* gm_nn_relu_bwd and gm_nn_bias_grad in one pass over nvox x C values:
 * g = (dA + dA2) where A > 0, else 0 (dA2, A nullable); dZ = g and dZb =
 * bf16(g) when given (both nullable); db[c] += sum over voxels of g[v][c].
 * C must be a multiple of 4 that divides 256; dA, dA2, A, dZ 16-byte and
 * dZb 8-byte aligned.  Sums in a different order from gm_nn_bias_grad. */
int gm_nn_relu_bias_bwd(int64_t nvox, int C, const float* dA, const float* dA2,
                        const float* A, float* dZ, uint16_t* dZb, float* db, void* stream);
/* loss[0] += 0.5 mean_v (Y[v][0] - target[v])^2 (FP64); dY (FP32, nvox x C)
 * = (Y[v][0] - target[v]) / nvox in channel 0, zero elsewhere. */
int gm_nn_mse(int64_t nvox, int C, const float* Y, const float* target, float* dY,
              double* loss, void* stream);

/* Energy-consistent loss of nb windows of nw = nx*ny*nz voxels each (x
 * fastest; Y: nb*nw*C FP32, channel 0 used; target: nb*nw): loss +=
 * 0.5 mean_v r_v^2 + 0.5 lambda mean_w m_w^2, with r = Y - target and m_w the
 * sum of r over the window's free voxels (not on the faces x = 0, x = max,
 * y = 0, y = max, z = 0) divided by nw: the window's heat-content error, which
 * the energy audit of a jump measures; dY = (r_v + lambda m_w [v free]) /
 * (nb nw) on channel 0, 0 on the others.  wsum: nb FP64 of device memory,
 * zeroed by the caller.  No reference counterpart (training loss of the
 * learned operator). */
int gm_nn_mse_energy(int64_t nw, int nb, int C, int nx, int ny, float lambda, const float* Y,
                     const float* target, double* wsum, float* dY, double* loss, void* stream);
/* Network input of one window of n = {nx, ny, nz} nodes (x fastest) at
 * geo = {x0, y0, z0, h}: 16 bf16 channels, 0: (T - T_amb)/T_scale, 1: the
 * Dirichlet face values (W, E, S, N, bottom) likewise, 0 inside, 2: the laser
 * intensity exp(-3((x-vx)^2 + (y-vy)^2)/r^2 - 3(z-vz)^2/d^2), 3: P/P_scale,
 * 4..15: 0.  prm = {T_amb, T_scale, vx, vy, vz, r, d, P, P_scale}. */
int gm_nn_window_input(const int32_t* n_host, const float* geo_host, const float* T,
                       const float* Tface, const float* prm_host, uint16_t* X, void* stream);
/* Tn = max(T_floor, T + T_scale * Y[v][0]) inside; faces take Tface (or T). */
int gm_nn_window_apply(const int32_t* n_host, const float* T, const float* Tface,
                       const float* Y, int C, float T_scale, float T_floor, float* Tn,
                       void* stream);
/* y += a x (FP32). */
int gm_nn_axpy(int64_t n, float a, const float* x, float* y, void* stream);

/* y[i] = x[i] for 0/1 phase bytes (the engine's uint8 S1) read as FP32 S1
 * (the form gm_state_properties takes, cF:2589).  No reference counterpart:
 * the reference keeps S1 as float throughout. */
int gm_u8_to_f32(int64_t n, const uint8_t* x, float* y, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GOMELT_H */
