// Internal types shared by the gfx950 kernels (gm_kernels.hip) and the host
// engine / C ABI (gm_engine.hip).  Nothing here is part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gomelt.h"

#define GM_MAX_ROWS 1024  // N2*N3 toolpath rows per L1 step
#define GM_MAX_N2 64

// Integer box of nodes (or elements): lo[d] .. lo[d]+n[d]-1.
struct GmBox {
    int32_t lo[3];
    int32_t n[3];
};

// Everything that changes from one L1 step to the next.  It lives in device
// memory and kernels read it through a pointer, so a recorded hipGraph stays
// valid while the windows move: the host only re-uploads this struct.
struct GmStepDesc {
    gm_grid g[4];          // current meshes; [1] L1, [2] L2, [3] L3
    gm_grid gprev[4];      // meshes before the last window move
    GmBox ov3;             // L2 nodes overlapped by L3 (L2 index space)
    gm_grid ov3_xyz;       // their coordinates (reference L3 overlapCoords)
    GmBox ov2;             // L1 nodes overlapped by L2 (L1 index space)
    gm_grid ov2_xyz;       // their coordinates (reference L2 overlapCoords)
    int32_t l0_lo3[3];     // Level0 node of L3 node (0,0,0); stride 1
    int32_t l0_lo2[3];     // Level0 node of L2 node (0,0,0)
    int32_t l0_st2[3];     // Level0 stride of L2 nodes
    int32_t l0_n[3];       // Level0 nodes per axis
    int64_t l1_ne_active, l1_nn_active;
    int64_t sub[4];        // substrate node counts per level
    int32_t nrows;
    float dt1;             // sum of row dt over the L1 step
    float dt2[GM_MAX_N2];  // sum of row dt per L2 substep
    float rows[GM_MAX_ROWS][7];
};

// Device copy of the material with derived constants (kernel argument).
#define GM_LUT_BINS 64
#define GM_PREC_BINS 32
struct DevMat {
    float T_amb, T_sol, T_liq, T_boil;
    float k_p, k_a0, k_a1, k_f;
    float cp_a0, cp_a1, cp_m, cp_f, rho;
    int32_t k_n, cp_n;
    float kT[GM_MAX_KNOTS], kv[GM_MAX_KNOTS], ks[GM_MAX_KNOTS];
    float cT[GM_MAX_KNOTS], cv[GM_MAX_KNOTS], cs[GM_MAX_KNOTS];
    float h_conv, sigma_eps, evc_cp, CT, CM, Lev, invTb, Tclip, Tamb4, sqCM;
    float eta, rsq, dsq, rc, dc, pc6;  // pc6 = 6*sqrt(3)*eta
    // k and cp on the union of the table knots (a property without a table
    // contributes its linear law): mseg[j] = (T_j, k(T_j), k slope, cp(T_j),
    // cp slope) on [T_j, T_j+1), linear extension at the ends; mn = 0 when
    // neither property has a table.  Segment of T: start at mlut[b] of the
    // bin b = (T - mT0) * miw, step up while T passed the next knot.
    int32_t mn, mnb;
    float mT0, miw;
    int32_t mlut[GM_LUT_BINS];
    float mseg[2 * GM_MAX_KNOTS][5];
    // Per-bin segment records (pnb > 0 when every bin of width 1/piw over
    // the knots, widened by a quarter bin on each side, meets at most two
    // segments A and A+1).  prec[b] = (T_{A+1} or +inf, T_A, k(T_A), k slope,
    // cp(T_A), cp slope, k(T_{A+1}), k slope, cp(T_{A+1}), cp slope, 0 x 6;
    // 64 bytes, so a record's offset is a shift):
    // the segment of T in bin b is A+1 when T >= prec[b][0], else A.
    int32_t pnb;
    float pT0, piw;
    float prec[GM_PREC_BINS][16];
};

// One axis of the tabulated restriction between a fine and a coarse mesh
// (the per-axis factors of the reference's Shapes tensors, cF:1213-1358).
// For fine element e and Gauss point g in {0,1} along the axis:
//   xq  : Gauss point coordinate
//   phi : coarse shape factors (x1 - xq)/hc, (xq - x0)/hc in the coarse
//         element cq that holds the Gauss point (clipped into the mesh)
//   c0  : coarse element holding Gauss point 0 (where the values are added)
// Fine elements with c0 == c form the contiguous range [ex_lo, ex_hi).
struct GmAxisTab {
    int32_t* c0;       // [nef]
    float4* q;         // [nef][2] : (xq, phi0, phi1, 0)
    int32_t* info;     // [0] c_lo, [1] c_n, [2 .. 2+cmax) ex_lo, then ex_hi
};

struct GmPair {
    int coarse, fine;  // level ids (grids read from the step descriptor)
    int nef[3];        // fine elements per axis
    int cmax[3];       // bound on covered coarse elements per axis
    GmAxisTab ax[3];
};

enum { GM_SRC_NONE = 0, GM_SRC_L3 = 1, GM_SRC_TABLES = 2 };
enum { GM_BC_NONE = 0, GM_BC_L1 = 1, GM_BC_FINE = 2 };
// dt selectors
enum { GM_DT_VALUE = -1, GM_DT_ROW = 0, GM_DT_L1 = 1, GM_DT_L2 = 2 };

// Arguments of the fused level step (solveMatrixFreeFE + computeStateProperties
// + sources + computeConvRadBC + Dirichlet + clamp [+ melt-time accumulation]).
struct LSArgs {
    int level;            // grid = desc->g[level] when use_desc
    int use_desc;
    gm_grid grid;         // used when !use_desc
    const float* T;
    const void* S1;       // float (s1_float) or uint8
    int s1_float;
    float* Tout;
    uint8_t* S1out;       // nullable
    const float* kin;     // optional precomputed nodal k, rho*cp
    const float* cin;
    int64_t n_sub;
    int64_t ne_active;    // elements 0..ne_active-1 take part
    int64_t nn_active;    // nodes >= nn_active are set to T_amb
    int dt_mode, dt_index;
    float dt;
    // source
    int src_mode;
    int src_row;          // GM_SRC_L3: desc row (or -1: src_v/src_P)
    float src_v[3], src_P;
    const float* G;       // GM_SRC_TABLES: [npos][3][G_nmax]
    int G_npos, G_nmax;
    float G_scale;
    const int32_t* G_info[3];
    const float* Fext;    // nullable full-size extra source
    int conv;             // top-face convection/radiation/evaporation
    // correction term (restricted T' terms), stored on a node box
    const float* corr;
    const int32_t* corr_info[3];
    int corr_nmax[3];
    const float* corr_full;  // nullable full-size correction
    // Dirichlet
    int bc_mode;
    float bc_vals[5];     // y0, y1, x0, x1, z0
    int bc_level;         // coarse level of the prolonged field
    const float* bcA;     // face values of the prolonged coarse fields (full size)
    const float* bcB;
    float alpha;
    int clamp;
    // melt-time accumulation (Part 2 of the L3 subcycle)
    int accum_on;
    const uint8_t* S2prev;
    uint8_t* S2out;
    float* accum;
    float* max_accum;
    // element operator of the level: kc[p] for neighbour patterns p and the
    // lumped mass factor (volume/8)
    float kc[8];
    float mrow;
    float w2, w3;         // face and volume Gauss weights (detJ)
    // nodal properties carried between steps: k[nn] then cp[nn] of T (in),
    // of Tout (out), in the raw units of the material tables
    const float* kcin;
    float* kcout;
    // tabulated separable L3 source factors of the desc rows: row r at
    // srcT + r * (nx + ny + nz): x factors (times 6 sqrt(3) eta P w3), then y, z
    const float* srcT;
    int prec_ok;          // the material has per-bin property records (pnb > 0)
    // correction read straight from the FP64 restriction accumulator (same
    // node box as corr), cleared as it is read, plus the coarse laser source
    // on the same box (corrG, times corrG_scale); replaces a finishing pass
    double* corr64;
    const float* corrG;
    float corrG_scale;
};

// One instance of a batched L3-type step: its step arguments (grid,
// buffers, coefficients, dt; source tables at a.srcT with a.src_row = 0),
// its material and laser constants, and its laser row of the step
// (x, y, z, Ljump, Ldwell, dt, P).
struct BatchItem {
    LSArgs a;
    const DevMat* m;
    float row[7];
    float w3;
};

// L3 face values of one engine (k_face_bc_batch): L2 fields A (new), B (old)
struct FaceItem {
    const GmStepDesc* d;
    const float* A;
    const float* B;
    float* fa;
    float* fb;
};

// T' and injection of one engine's L3 window into its L2 field (k_tprime_batch)
struct TpItem {
    const GmStepDesc* d;
    const float* fT;
    float* cT;
    float* Tp;
    int ntp;
    const DevMat* m;
    const uint8_t* cS1;
    int64_t c_nsub;
    float* cKC;
};

// Restriction (fine -> coarse) arguments.
struct RArgs {
    const float* T;       // fine temperature (for element k and rho*cp)
    const uint8_t* S1;    // fine S1 (uint8)
    const float* S1f;     // or FP32 S1
    int64_t n_sub;
    const float* Tp;      // flux T' (nullable)
    const float* TpA;     // mass dT' = TpA - TpB (nullable)
    const float* TpB;
    float dt;
    int dt_mode, dt_index;
    float wf;             // fine volume Gauss weight
    // node-centric path (fine nodes on every coarse node): the fine element
    // operators and, optionally, the carried nodal k, cp of T
    const float* KC;      // k[nn] then cp[nn] (nullable: computed from T, S1)
    float kcs[8];         // stiffness per neighbour pattern (unit k)
    float mcs[8];         // consistent mass per neighbour pattern (unit rho*cp)
    // wave-brick kernel (carried k, cp): the x ratio and the first fine x on a
    // coarse node, fixed while the pair's meshes do not move relative to each
    // other by a fraction of a coarse spacing (host-checked)
    int dpp_ok, R0, a0;
    int R1, R2, W;        // y, z ratios; strip width R0 * floor(61 / R0)
    double invR[3];       // 1 / ratio per axis (FP64 weights)
};

// One engine's restriction in a batched launch (launch_restrict_batch): the
// arguments of its single launch.
struct RItem {
    GmPair p;
    const GmStepDesc* d;
    RArgs r;
    double* acc;
};

// ---- launchers (gm_kernels.hip) -------------------------------------------
namespace gm {

// Tile index of block b of nb, mapped so that each XCD (blocks are dealt to
// the 8 XCDs round-robin) works on a contiguous run of tiles and shares halos
// through its own L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    const int q = nb >> 3, r = nb & 7;
    const int x = b & 7, k = b >> 3;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// 3x3x3 BF16 MFMA convolutions of the learned operator (gm_conv.hip);
// shape = {nb, D, H, W, cin, cout}
bool conv_shape_ok(const int32_t* shape);
void launch_conv_fwd(const int32_t* shape, const uint16_t* X, const uint16_t* Wk,
                     const float* bias, int relu, float* Y, uint16_t* Ybf, hipStream_t s);
// bytes of the K-split workspace the forward would use for this shape (0: no split)
int64_t conv_fwd_split_bytes(const int32_t* shape);
void launch_conv_fwd_ws(const int32_t* shape, const uint16_t* X, const uint16_t* Wk,
                        const float* bias, int relu, float* Y, uint16_t* Ybf, void* ws,
                        int64_t ws_bytes, hipStream_t s);
void launch_conv_wflip(int ci, int co, const uint16_t* Wk, uint16_t* Wt, hipStream_t s);
void launch_conv_wgrad(const int32_t* shape, const uint16_t* X, const uint16_t* dY, int vsplit,
                       float* dW, hipStream_t s);
void launch_adam(float* w, uint16_t* wbf, const float* g, float* m, float* v, int64_t n, float lr,
                 float b1, float b2, float eps, int step, hipStream_t s);
void launch_f32_to_bf16(const float* a, uint16_t* b, int64_t n, hipStream_t s);

// elementwise and reduction kernels of the operator's network (gm_nn.hip)
void launch_nn_pool2(int dir, int nb, int D, int H, int W, int C, const void* in, void* out,
                     hipStream_t s);
void launch_nn_up2(int dir, int nb, int D, int H, int W, int C, const void* in, void* out,
                   hipStream_t s);
void launch_nn_concat(int64_t nvox, int ca, int cb, const uint16_t* A, const uint16_t* B,
                      uint16_t* Y, hipStream_t s);
void launch_nn_split(int64_t nvox, int ca, int cb, const float* dY, float* dA, float* dB,
                     hipStream_t s);
void launch_nn_relu_bwd(int64_t n, const float* dA, const float* dA2, const float* A, float* dZ,
                        uint16_t* dZb, hipStream_t s);
void launch_nn_bias_grad(int64_t nvox, int C, const float* dZ, float* db, hipStream_t s);
void launch_nn_relu_bias_bwd(int64_t nvox, int C, const float* dA, const float* dA2,
                             const float* A, float* dZ, uint16_t* dZb, float* db, hipStream_t s);
void launch_nn_mse(int64_t nvox, int C, const float* Y, const float* target, float* dY,
                   double* loss, hipStream_t s);
void launch_nn_window_input(const int32_t* n, const float* geo, const float* T, const float* Tface,
                            const float* prm, uint16_t* X, hipStream_t s);
void launch_nn_axpy(int64_t n, float a, const float* x, float* y, hipStream_t s);
void launch_nn_u8_to_f32(int64_t n, const uint8_t* x, float* y, hipStream_t s);
void launch_nn_mse_energy(int64_t nw, int nb, int C, int nx, int ny, float lambda,
                          const float* Y, const float* target, double* wsum, float* dY,
                          double* loss, hipStream_t s);
void launch_nn_window_apply(const int32_t* n, const float* T, const float* Tface, const float* Y,
                            int C, float T_scale, float T_floor, float* Tn, hipStream_t s);

// batched L3-type steps
#define GM_BATCH_TILES 5
extern const int kBatchTiles[GM_BATCH_TILES][2];  // {rows per band, planes per brick}
int batch_bricks(const gm_grid& g, int yb, int zc);
void launch_batch_src(const BatchItem* it, int n, int maxaxes, hipStream_t s);
void launch_batch_step(const BatchItem* it, const DevMat* mats, const int* bstart, int n,
                       int nblocks, int tile, bool rec, const GmStepDesc* d, hipStream_t s);
void launch_batch_step_acc(const BatchItem* it, const DevMat* mats, const int* bstart, int n,
                           int nblocks, int tile, bool rec, const GmStepDesc* d, hipStream_t s);
// batched L2 steps (bricks 1 x 2), face values and T' + injection of many engines
void launch_batch_step_l2(const BatchItem* it, const DevMat* mats, const int* bstart, int n,
                          int nblocks, bool rec, const GmStepDesc* d, hipStream_t s);
void launch_face_bc_batch(const FaceItem* it, const int* bstart, int n, int nblocks,
                          hipStream_t s);
int face_bc_blocks(const gm_grid& fine_grid);
void launch_tprime_batch(const TpItem* it, const int* bstart, int n, int nblocks,
                         hipStream_t s);
// the node-aligned restriction with carried k, cp (the form the batched
// launch runs), its blocks, and the batched launch
bool restrict_batchable(const RArgs& r);
int restrict_batch_blocks(const GmPair& p, const RArgs& r);
void launch_restrict_batch(const RItem* it, const int* bstart, int n, int nblocks, bool flux,
                           bool mass, hipStream_t s);
// blocks of one item in launch_tprime_batch; *ntp_out: its T' blocks of the single launch
int tprime_blocks(const gm_grid& fine_grid, int64_t nbox, int* ntp_out);
void launch_step_tile(const LSArgs& a, const GmStepDesc* d, const gm_grid& hg, const DevMat* m,
                      int tile, bool rec, hipStream_t s);
DevMat make_devmat(const gm_material& m);
void level_coeffs(const gm_grid& g, float kc[8], float* mrow, float* w2, float* w3);

void launch_level_step(const LSArgs& a, const GmStepDesc* d, const gm_grid& host_grid,
                       const DevMat* m_dev, hipStream_t s);
// fine-level step that reads k, cp of its input and writes those of its output
void launch_level_step_kc(const LSArgs& a, const GmStepDesc* d, const gm_grid& host_grid,
                          const DevMat* m_dev, hipStream_t s);
// k, cp of (T, S1) for a level step that reads them
// separable source factors of the L3 mesh for desc rows [0, nrows)
void launch_src3_tables(const GmStepDesc* d, const DevMat* m_dev, int nrows, int ntot, float w3,
                        float* out, hipStream_t s);
void launch_acc_l3(const GmStepDesc* d, float* accum, float* max_accum, float* acc3, float* macc3,
                   int64_t n3, int to_l3, hipStream_t s);
void launch_props_kc(const DevMat* m_dev, const float* T, const uint8_t* S1, int64_t nn,
                     int64_t n_sub, float* kc, hipStream_t s, const float* T2 = nullptr,
                     const uint8_t* S12 = nullptr, int64_t nn2 = 0, int64_t n_sub2 = 0,
                     float* kc2 = nullptr);
// device-to-device copies batched into one launch (hipMalloc'd buffers)
struct CopyItem {
    void* dst;
    const void* src;
    int64_t bytes;
};
struct CopyList {
    CopyItem c[8];
};
void launch_copy_multi(const CopyList& L, int n, hipStream_t s);
void launch_axis_tables(const GmPair* p, int np, const GmStepDesc* d, hipStream_t s);
void launch_source_tables(const GmPair& p, const GmStepDesc* d, const DevMat& m,
                          int row0, int npos, int use_rows, const float* rows_dev,
                          float* G, int nmax, hipStream_t s);
void launch_restrict(const GmPair& p, const GmStepDesc* d, const DevMat& m, const RArgs& r,
                     double* partial, hipStream_t s);
// node-centric restriction; adds into the FP64 node box acc of the pair
void launch_restrict_fast(const GmPair& p, const GmStepDesc* d, const DevMat* m, const RArgs& r,
                          double* acc, hipStream_t s);
// V = Vbase + sum of the pairs' FP64 node boxes (+ coarse source); the node
// boxes are cleared for the next accumulation
void launch_restrict_nodes_acc(const GmPair& outer, double* acc_outer, const GmPair* inner,
                               double* acc_inner, const float* Vbase, float* V,
                               const GmPair* gpair, const float* F, float scale, hipStream_t s);
// coarse source of nk position groups (npos each) on the node box of pair p
void launch_src_box(const GmPair& p, const float* G, int nk, int npos, int nmax, float* F,
                    hipStream_t s);
void launch_restrict_nodes(const GmPair& outer, const double* part_outer, const GmPair* inner,
                           const double* part_inner, const float* Vbase, float* V,
                           const GmPair* gpair, const float* G, int npos, int nmax, float scale,
                           hipStream_t s);
// (with cKC: also k, cp of the injected coarse nodes, from cS1)
void launch_tprime_inject(const GmStepDesc* d, int fine_level, int coarse_level,
                          const float* fineT, float* coarseT, float* Tp, const gm_grid& fine_grid,
                          int64_t nbox, const DevMat* m, const uint8_t* cS1, int64_t c_nsub,
                          float* cKC, hipStream_t s);
void launch_face_bc(const GmStepDesc* d, int fine, int coarse, const gm_grid& gf, const float* A,
                    const float* B, float* fa, float* fb, hipStream_t s);
void launch_interp_grid(const GmStepDesc* d, int src_level, const gm_grid* src_host,
                        const float* u, const uint8_t* u8, int tgt_mode, const gm_grid* tgt_host,
                        int64_t n_target, float* out, hipStream_t s);
void launch_inject(const GmStepDesc* d, int fine_level, int coarse_level, const float* fineT,
                   float* coarseT, int64_t n_box, hipStream_t s);
void launch_tprime(const GmStepDesc* d, int fine_level, int coarse_level, const float* fineT,
                   const float* coarseT, float* Tp, int64_t n, hipStream_t s);
void launch_move_l3(const GmStepDesc* d, const float* T1, const float* Tp2, const float* Tp3,
                    float* T3n, float* Tp3n, int64_t nn3, hipStream_t s);
void launch_move_l2(const GmStepDesc* d, const float* T1, const float* Tp2, float* T2n,
                    float* Tp2n, int64_t nn2, hipStream_t s);
void launch_l1_s1_from_l2(const GmStepDesc* d, const uint8_t* S1_2, float* S1_1, int64_t n_box,
                          hipStream_t s);
void launch_fill_f32(float* p, float v, int64_t n, hipStream_t s);
void launch_fill_u8(uint8_t* p, uint8_t v, int64_t n, hipStream_t s);
void launch_copy_f32(float* dst, const float* src, int64_t n, hipStream_t s);
void launch_copy_u8(uint8_t* dst, const uint8_t* src, int64_t n, hipStream_t s);
void launch_clamp(float* T, int64_t n, float Tamb, hipStream_t s);
void launch_l0_gather(const GmStepDesc* d, const uint8_t* L0S1, const uint8_t* s2snap,
                      const int32_t* s2box, uint8_t* S1_3, uint8_t* S2_3, uint8_t* S1_2,
                      int64_t nn3, int64_t nn2, hipStream_t s);
void launch_l0_scatter(const GmStepDesc* d, const uint8_t* S1_3, const uint8_t* S2_3,
                       uint8_t* L0S1, uint8_t* s2snap, int32_t* s2box, int64_t nn3,
                       hipStream_t s);
void launch_melt_single(const GmStepDesc* d, const float* T3, const uint8_t* preS2,
                        const uint8_t* S2, float dt, float Tliq, float* accum, float* max_accum,
                        int64_t nn3, hipStream_t s);
void launch_minmax(const float* T, int64_t n, float* out2, hipStream_t s);
// melt-pool size of a level's field into dims (8 floats of device memory)
void launch_melt_pool(const gm_grid& g, const float* T, float Tl, float* dims, hipStream_t s);
// energy audit of an explicit update of one level (FP64 sums, 8 doubles)
struct EnergyCoef {
    double kc[8];   // unit-k box-element stiffness by neighbour pattern
    double vol8;    // element volume / 8
};
EnergyCoef energy_coef(const gm_grid& g);
void launch_energy(const gm_grid& g, int64_t ne_active, const float* k, const float* rc,
                   const float* T, const float* F, const float* Corr, const float* Tn,
                   int bc_faces, float T_clamp, float dt, double* sums, hipStream_t s);
void launch_props(const DevMat& m, const float* T, const float* S1, int64_t nn, int64_t nsub,
                  float* S1o, uint8_t* S2o, float* k, float* c, hipStream_t s);
void launch_source_l3_full(const gm_grid& g, const DevMat& m, float vx, float vy, float vz,
                           float P, float* F, hipStream_t s);
void launch_source_l3_mean(const gm_grid& g, const DevMat& m, const float* rows, int n_rows,
                           float* F, hipStream_t s);
void launch_convrad_full(const gm_grid& g, const DevMat& m, const float* T, int64_t ne_active,
                         float* F, hipStream_t s);
void launch_interp_matrix(const gm_grid& src, const gm_grid& tgt, float* Nc, int32_t* node,
                          hipStream_t s);
void launch_interp_w(const float* Nc, const int32_t* node, int64_t n, const float* T,
                     float* out, hipStream_t s);
void launch_shift_l0(uint8_t* L0S1, float* accum, int64_t n1, int64_t n2, int64_t nn,
                     hipStream_t s);
}  // namespace gm
