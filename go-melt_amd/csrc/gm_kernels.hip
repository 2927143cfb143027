// gfx950 kernels of the multilevel explicit FE heat step.
//
// Layout: every level field is a flat FP32 (or uint8) array in the x-fastest
// node numbering of the reference (i = ix + nx*(iy + ny*iz), cF:645-689).
// The hot kernel is k_level_step: one launch advances one level by one time
// step, fusing the reference's computeStateProperties (cF:2567-2614), the
// matrix-free element operator solveMatrixFreeFE (cF:582-642), the laser
// source (cF:2960-3012 / 2667-2730), the top-face flux (cF:2207-2301), the
// Dirichlet overwrite (cF:1568-1620), the inactive-node reset (cF:1848-1865)
// and the T_amb clamp.  It is node-centric and marches along z: a block owns
// 32x8 node columns and 4 output planes, keeps three T planes and two element
// planes (k, rho*cp means) in LDS, evaluates the node properties once per
// loaded plane node, and applies the 27-point stencil the box element's 8x8
// matrices reduce to, summed by neighbour class.  No atomics, deterministic.
#include <climits>
#include <math.h>
#include <algorithm>
#include <vector>

#include "gm_internal.h"

namespace gm {

// 1-D Gauss factors at xi = -/+ 1/sqrt(3): N1 same side / other side.
#define GA 0.78867513459481287f
#define GB 0.21132486540518713f

__device__ __forceinline__ float coordf(const gm_grid& g, int d, int i) {
    return g.x0[d] + (float)i * g.h[d];
}

__device__ __forceinline__ int64_t nid(int ix, int iy, int iz, int nx, int ny) {
    return (int64_t)ix + (int64_t)nx * ((int64_t)iy + (int64_t)ny * iz);
}

// Piecewise-linear table with linear extension beyond the end knots.
__device__ __forceinline__ float tab_eval(const float* kT, const float* kv, const float* ks, int n,
                                          float T) {
    float b = kv[0], t0 = kT[0], sl = ks[0];
#pragma unroll
    for (int m = 1; m < GM_MAX_KNOTS - 1; ++m) {
        if (m < n - 1 && T >= kT[m]) {
            b = kv[m];
            t0 = kT[m];
            sl = ks[m];
        }
    }
    return b + sl * (T - t0);
}

// computeStateProperties (cF:2584-2612) for one node.  With S1, S2, S3 in
// {0,1} the reference's sums of products are exact selects.
__device__ __forceinline__ void props(const DevMat& m, float T, float s1in, bool substrate,
                                      float& k, float& c, bool& s1, bool& s2) {
    s2 = T >= m.T_liq;
    const bool s3 = (T > m.T_sol) && (T < m.T_liq);
    s1 = (s1in > 0.499f) || s2 || substrate;
    const float kb = m.k_n ? tab_eval(m.kT, m.kv, m.ks, m.k_n, T) : (m.k_a1 * T + m.k_a0);
    const float kk = s2 ? m.k_f : (s1 ? kb : m.k_p);
    k = kk / 1000.0f;
    const float cs = m.cp_n ? tab_eval(m.cT, m.cv, m.cs, m.cp_n, T) : (m.cp_a1 * T + m.cp_a0);
    const float cp = s2 ? m.cp_f : (s3 ? m.cp_m : cs);
    c = m.rho * cp;
}

// interpolatePointsMatrix rule (cF:1060-1104): coarse element clipped into
// the mesh, trilinear weights, all zero if any leaves [-0.01, 1.01].
__device__ __forceinline__ void interp8(const gm_grid& c, float x, float y, float z, float N[8],
                                        int64_t nd[8]) {
    int e[3];
    const float q[3] = {x, y, z};
    float lo[3], hi[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        int ee = (int)floorf((q[d] - c.x0[d]) / c.h[d]);
        ee = ee < 0 ? 0 : (ee > c.n[d] - 2 ? c.n[d] - 2 : ee);
        e[d] = ee;
        const float x0 = coordf(c, d, ee), x1 = coordf(c, d, ee + 1);
        lo[d] = x1 - q[d];
        hi[d] = q[d] - x0;
    }
    const float inv = 1.0f / (c.h[0] * c.h[1] * c.h[2]);
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const int bx = ((a + 1) >> 1) & 1, by = (a >> 1) & 1, bz = a >> 2;
        const float v = (bx ? hi[0] : lo[0]) * (by ? hi[1] : lo[1]) * (bz ? hi[2] : lo[2]) * inv;
        N[a] = v;
        ok = ok && (v >= -1e-2f) && (v <= 1.0f + 1e-2f);
        nd[a] = nid(e[0] + bx, e[1] + by, e[2] + bz, c.n[0], c.n[1]);
    }
#pragma unroll
    for (int a = 0; a < 8; ++a) N[a] = ok ? fminf(fmaxf(N[a], 0.0f), 1.0f) : 0.0f;
}

// Node index -> (ix, iy, iz) in 32-bit unsigned arithmetic (every mesh and
// box here has fewer than 2^32 nodes; 64-bit division costs ~10x more).
__device__ __forceinline__ void dec3(int64_t t, int nx, int ny, int& ix, int& iy, int& iz) {
    uint32_t u = (uint32_t)t;
    const uint32_t q = u / (uint32_t)nx;
    ix = (int)(u - q * (uint32_t)nx);
    iz = (int)(q / (uint32_t)ny);
    iy = (int)(q - (uint32_t)iz * (uint32_t)ny);
}

// interp8 returning the clipped coarse element per axis instead of node ids
__device__ __forceinline__ void interp8e(const gm_grid& c, float x, float y, float z, float N[8],
                                         int e[3]) {
    const float q[3] = {x, y, z};
    float lo[3], hi[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        int ee = (int)floorf((q[d] - c.x0[d]) / c.h[d]);
        ee = ee < 0 ? 0 : (ee > c.n[d] - 2 ? c.n[d] - 2 : ee);
        e[d] = ee;
        const float x0 = coordf(c, d, ee), x1 = coordf(c, d, ee + 1);
        lo[d] = x1 - q[d];
        hi[d] = q[d] - x0;
    }
    const float inv = 1.0f / (c.h[0] * c.h[1] * c.h[2]);
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const int bx = ((a + 1) >> 1) & 1, by = (a >> 1) & 1, bz = a >> 2;
        const float v = (bx ? hi[0] : lo[0]) * (by ? hi[1] : lo[1]) * (bz ? hi[2] : lo[2]) * inv;
        N[a] = v;
        ok = ok && (v >= -1e-2f) && (v <= 1.0f + 1e-2f);
    }
#pragma unroll
    for (int a = 0; a < 8; ++a) N[a] = ok ? fminf(fmaxf(N[a], 0.0f), 1.0f) : 0.0f;
}

__device__ __forceinline__ float interp_f(const gm_grid& c, const float* u, float x, float y,
                                          float z) {
    float N[8];
    int64_t nd[8];
    interp8(c, x, y, z, N, nd);
    float s = 0.0f;
#pragma unroll
    for (int a = 0; a < 8; ++a) s += N[a] * u[nd[a]];
    return s;
}

// Top-face flux density (cF:2269-2290), W/mm^2.
__device__ __forceinline__ float face_flux(const DevMat& m, float Tq) {
    Tq = (Tq > m.Tclip) ? m.Tclip : Tq;
    const float invT = 1.0f / Tq;
    const float E = m.Lev + m.cp_f * (Tq - m.T_amb);
    const float mol = sqrtf(m.CM * invT);
    const float S = m.evc_cp * expf(-m.CT * (invT - m.invTb)) * mol * E;
    const float t2 = Tq * Tq;
    const float q = m.h_conv * (m.T_amb - Tq) + m.sigma_eps * (m.Tamb4 - t2 * t2) - S;
    return q * 1e-6f;
}

// 2x2 face Gauss point / node shape values (cF:778-815).
__device__ __forceinline__ float n2(int q, int a) {
    const float s = 0.62200846792814621f, j = 0.16666666666666667f, o = 0.04465819873852045f;
    const int qx = ((q + 1) >> 1) & 1, qy = (q >> 1) & 1;
    const int ax = ((a + 1) >> 1) & 1, ay = (a >> 1) & 1;
    const int dif = (qx != ax) + (qy != ay);
    return dif == 0 ? s : (dif == 1 ? j : o);
}

// Separable laser source factor along one axis at node i (cF:2986-3010 with
// the Gaussian of cF:1013-1025): sum over the elements on either side and
// their two Gauss points of N1 * coef * exp(-3 (xq - v)^2 / sq).
__device__ __forceinline__ float src_axis(const gm_grid& g, int d, int i, float v, float sq,
                                          float coef) {
    float s = 0.0f;
    const int ne = g.n[d] - 1;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int e = i - 1 + k;
        if (e < 0 || e >= ne) continue;
        const float xa = coordf(g, d, e), xb = coordf(g, d, e + 1);
        const int a = i - e;
#pragma unroll
        for (int gg = 0; gg < 2; ++gg) {
            const float xq = gg == 0 ? (xa * GA + xb * GB) : (xa * GB + xb * GA);
            const float t = xq - v;
            const float qv = coef * expf((-3.0f * (t * t)) / sq);
            s += (a == gg ? GA : GB) * qv;
        }
    }
    return s;
}

// floor(u / R) for 0 <= u < 2^20 from a reciprocal of R (rcp error far below
// the 0.5 / R distance of (u + 0.5) / R from an integer)
__device__ __forceinline__ int fdiv_small(int u, float invR) {
    return (int)(((float)u + 0.5f) * invR);
}

// Workgroup barrier for LDS hand-offs only.  __syncthreads() also waits for
// every outstanding global load and store of the wave (vmcnt(0)), which in a
// plane march puts the store round trip of each plane on the critical path.
// Here a wave waits only for its own LDS operations (lgkmcnt(0)) before
// s_barrier; the empty asm statements keep the compiler from moving memory
// accesses across it.
__device__ __forceinline__ void lds_barrier() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xc07f);  // vmcnt and expcnt at their maxima, lgkmcnt(0)
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// Tiles of LX x LY node columns; ZCT output planes per block (template).
#define LX 32
#define LY 8
#define PX (LX + 2)
#define PY (LY + 2)
#define PN (PX * PY)
#define QX (LX + 1)
#define QY (LY + 1)
#define QN (QX * QY)
// LDS words of the staged property table: the per-bin segment records, or
// the loop lookup (bin -> first segment, then the segment rows).
#define PTW (GM_PREC_BINS * 16 > GM_LUT_BINS + 10 * GM_MAX_KNOTS ? GM_PREC_BINS * 16 \
                                                                 : GM_LUT_BINS + 10 * GM_MAX_KNOTS)

// computeStateProperties (cF:2584-2612) for one node, without the unit
// factors: k in W/(m K), cp in J/(kg K) (the 1/1000 of cF:2602 and rho of
// cF:2612 are folded into the host-side stencil and mass constants).  With
// S1, S2, S3 in {0,1} the reference's sums of products are exact selects.
struct PropC {
    float Tliq, Tsol, kp, kf, cpf, cpm;
    int mn, mnb;   // loop lookup (used when pnb == 0)
    float mT0, miw;
    int pnb;       // per-bin records
    float pT0, piw;
};

__device__ __forceinline__ PropC load_pc(const DevMat* __restrict__ m) {
    PropC pc;
    pc.Tliq = m->T_liq;
    pc.Tsol = m->T_sol;
    pc.kp = m->k_p;
    pc.kf = m->k_f;
    pc.cpf = m->cp_f;
    pc.cpm = m->cp_m;
    pc.mn = m->mn;
    pc.mnb = m->mnb;
    pc.mT0 = m->mT0;
    pc.miw = m->miw;
    pc.pnb = m->pnb;
    pc.pT0 = m->pT0;
    pc.piw = m->piw;
    return pc;
}

// Property table into LDS (PTW floats): the records when the material has
// them, else the loop lookup.
__device__ __forceinline__ void stage_tables(const DevMat* __restrict__ m, float* sTab, int tid,
                                             int nthr) {
    if (m->pnb > 0) {
        const float4* src = (const float4*)&m->prec[0][0];
        float4* dst = (float4*)sTab;
        for (int l = tid; l < GM_PREC_BINS * 4; l += nthr) dst[l] = src[l];
        return;
    }
    const int n = GM_LUT_BINS + 5 * m->mn;
    for (int l = tid; l < n; l += nthr)
        sTab[l] = l < GM_LUT_BINS ? __int_as_float(m->mlut[l])
                                  : m->mseg[(l - GM_LUT_BINS) / 5][(l - GM_LUT_BINS) % 5];
}

__device__ __forceinline__ void phase_select(const PropC& c, float T, bool s1in, bool substrate,
                                             float kb, float cs, float& kk, float& cp, bool& s1) {
    const bool s2 = T >= c.Tliq;
    const bool s3 = (T > c.Tsol) && (T < c.Tliq);
    s1 = s1in || s2 || substrate;
    kk = s2 ? c.kf : (s1 ? kb : c.kp);
    cp = s2 ? c.cpf : (s3 ? c.cpm : cs);
}

// Record lookup: one bin read, one compare, no loop.
__device__ __forceinline__ void props_rec(const PropC& c, const float* __restrict__ tab, float T,
                                          bool s1in, bool substrate, float& kk, float& cp,
                                          bool& s1) {
    int b = (int)((T - c.pT0) * c.piw);
    b = b < 0 ? 0 : (b > c.pnb - 1 ? c.pnb - 1 : b);
    const float4* r = (const float4*)(tab + 16 * b);  // 64-byte records: a shift
    const float4 r0 = r[0], r1 = r[1], r2 = r[2];
    const bool up = T >= r0.x;
    const float Ts = up ? r0.x : r0.y;
    const float k0 = up ? r1.z : r0.z, ks = up ? r1.w : r0.w;
    const float c0 = up ? r2.x : r1.x, cs = up ? r2.y : r1.y;
    const float dT = T - Ts;
    phase_select(c, T, s1in, substrate, k0 + ks * dT, c0 + cs * dT, kk, cp, s1);
}

__device__ __forceinline__ void props_c(const PropC& c, const float* __restrict__ tab, float T,
                                        bool s1in, bool substrate, float& kk, float& cp, bool& s1) {
    if (c.pnb > 0) {
        props_rec(c, tab, T, s1in, substrate, kk, cp, s1);
        return;
    }
    int b = (int)((T - c.mT0) * c.miw);
    b = b < 0 ? 0 : (b > c.mnb - 1 ? c.mnb - 1 : b);
    int j = __float_as_int(tab[b]);
    const float* sg = tab + GM_LUT_BINS;
    while (j + 1 < c.mn - 1 && T >= sg[5 * (j + 1)]) ++j;
    const float dT = T - sg[5 * j];
    phase_select(c, T, s1in, substrate, sg[5 * j + 1] + sg[5 * j + 2] * dT,
                 sg[5 * j + 3] + sg[5 * j + 4] * dT, kk, cp, s1);
}

// Modes fixed at compile time: S1 storage, laser source, Dirichlet kind,
// melt-time accumulation, given k / rho*cp arrays, planes per block.
template <bool S1F, bool SRC, int BC, bool ACC, bool KIN, int ZCT>
__global__ void __launch_bounds__(256) k_level_step(LSArgs a, const GmStepDesc* __restrict__ d,
                                                    const DevMat* __restrict__ m) {
    constexpr int NP = ZCT + 2;              // planes loaded per block
    __shared__ float sT[3][PN];              // T of node planes (ring)
    __shared__ float sKn[PN], sCn[PN];       // k, cp of the plane being loaded
    __shared__ float sSk[2][QN], sSc[2][QN];  // x-y quad sums of k, cp (ring)
    __shared__ float sKe[2][QN], sCe[2][QN];  // element planes below / above (ring)
    __shared__ float sSrc[LX + LY + ZCT];
    __shared__ float sTab[PTW] __attribute__((aligned(16)));

    const gm_grid g = a.use_desc ? d->g[a.level] : a.grid;
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const int nxy = nx * ny;
    const int ntx = (nx + LX - 1) / LX, nty = (ny + LY - 1) / LY;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int ox = (t % ntx) * LX, oy = ((t / ntx) % nty) * LY, oz = (t / (ntx * nty)) * ZCT;
    const int nex = nx - 1, ney = ny - 1;
    const int nez_act = (int)(a.ne_active / ((int64_t)nex * ney));
    const int tid = threadIdx.x;
    const int tx = tid % LX, ty = tid / LX;
    const int ix = ox + tx, iy = oy + ty;
    const bool col_ok = ix < nx && iy < ny;

    PropC pc;
    if (!KIN) pc = load_pc(m);
    const float Tamb = m->T_amb;

    // ---- prologue: every global load of the block is issued here -----------
    // plane entries of this thread: l = tid and tid + 256
    int eoff[2];
    bool eok[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const int l = tid + 256 * e;
        const int jx = ox + l % PX - 1, jy = oy + l / PX - 1;
        eok[e] = l < PN && jx >= 0 && jx < nx && jy >= 0 && jy < ny;
        eoff[e] = jx + nx * jy;
    }
    float rT[NP][2], rS[NP][2], rK[NP][2];
#pragma unroll
    for (int pl = 0; pl < NP; ++pl) {
        const int p = oz - 1 + pl;
        const bool pin = p >= 0 && p < nz;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            rT[pl][e] = 0.0f;
            rS[pl][e] = 0.0f;
            rK[pl][e] = 0.0f;
            if (pin && eok[e]) {
                const int idx = eoff[e] + nxy * p;
                rT[pl][e] = a.T[idx];
                if (KIN) {
                    rK[pl][e] = a.kin[idx];
                    rS[pl][e] = a.cin[idx];
                } else {
                    rS[pl][e] = S1F ? ((const float*)a.S1)[idx]
                                    : (float)((const uint8_t*)a.S1)[idx];
                }
            }
        }
    }
    // per output node: correction (+ extra source), Dirichlet face values,
    // melt-time state
    const int cidx = ix + nx * iy;
    const bool face_col = ix == 0 || ix == nx - 1 || iy == 0 || iy == ny - 1;
    float rC[ZCT], rFa[ZCT], rFb[ZCT], rAcc[ZCT], rMax[ZCT];
    bool rS2[ZCT];
    int clo0 = 0, clo1 = 0, clo2 = 0, cn0 = 0, cn1 = 0, cn2 = 0;
    if (a.corr) {
        clo0 = a.corr_info[0][0]; cn0 = a.corr_info[0][1] + 1;
        clo1 = a.corr_info[1][0]; cn1 = a.corr_info[1][1] + 1;
        clo2 = a.corr_info[2][0]; cn2 = a.corr_info[2][1] + 1;
    }
    int l0base = 0;
    bool l0xy = false;
    if (ACC) {
        const int lx0 = d->l0_lo3[0] + ix, ly0 = d->l0_lo3[1] + iy;
        l0xy = lx0 >= 0 && lx0 < d->l0_n[0] && ly0 >= 0 && ly0 < d->l0_n[1];
        l0base = lx0 + d->l0_n[0] * ly0;
    }
#pragma unroll
    for (int z = 0; z < ZCT; ++z) {
        const int zo = oz + z;
        const bool ok = col_ok && zo < nz;
        const int idx = cidx + nxy * zo;
        float c = 0.0f;
        if (ok && a.corr) {
            const int bx = ix - clo0, by = iy - clo1, bz = zo - clo2;
            if (bx >= 0 && bx < cn0 && by >= 0 && by < cn1 && bz >= 0 && bz < cn2)
                c = a.corr[bx + a.corr_nmax[0] * (by + a.corr_nmax[1] * bz)];
        }
        if (ok && a.corr_full) c += a.corr_full[idx];
        if (ok && a.Fext) c += a.Fext[idx];
        rC[z] = c;
        rFa[z] = rFb[z] = 0.0f;
        if (BC == GM_BC_FINE && ok && (face_col || zo == 0)) {
            rFa[z] = a.bcA[idx];
            rFb[z] = a.bcB[idx];
        }
        rAcc[z] = rMax[z] = 0.0f;
        rS2[z] = false;
        if (ACC && ok) {
            rS2[z] = a.S2prev[idx] != 0;
            const int lz0 = d->l0_lo3[2] + zo;
            if (l0xy && lz0 >= 0 && lz0 < d->l0_n[2]) {
                const int64_t l0 = (int64_t)l0base + (int64_t)d->l0_n[0] * d->l0_n[1] * lz0;
                rAcc[z] = a.accum[l0];
                rMax[z] = a.max_accum[l0];
            }
        }
    }

    float dt = a.dt;
    if (a.dt_mode == GM_DT_ROW) dt = d->rows[a.dt_index][5];
    else if (a.dt_mode == GM_DT_L1) dt = d->dt1;
    else if (a.dt_mode == GM_DT_L2) dt = d->dt2[a.dt_index];
    const float mfac = a.mrow / dt;

    // per-block constants: table lookups and the separable source
    if (!KIN) {
        stage_tables(m, sTab, tid, 256);
    }
    if (SRC) {
        float v0, v1, v2, P;
        if (a.src_row >= 0) {
            v0 = d->rows[a.src_row][0];
            v1 = d->rows[a.src_row][1];
            v2 = d->rows[a.src_row][2];
            P = d->rows[a.src_row][6];
        } else {
            v0 = a.src_v[0];
            v1 = a.src_v[1];
            v2 = a.src_v[2];
            P = a.src_P;
        }
        if (tid < LX) {
            const int i = ox + tid;
            sSrc[tid] = i < nx ? src_axis(g, 0, i, v0, m->rsq, m->rc) * (m->pc6 * P) : 0.0f;
        } else if (tid < LX + LY) {
            const int i = oy + tid - LX;
            sSrc[tid] = i < ny ? src_axis(g, 1, i, v1, m->rsq, m->rc) : 0.0f;
        } else if (tid < LX + LY + ZCT) {
            const int i = oz + tid - LX - LY;
            sSrc[tid] = i < nz ? src_axis(g, 2, i, v2, m->dsq, m->dc) : 0.0f;
        }
    }
    // quad-sum entries of this thread: q = tid and tid + 256
    int qh[2], qo[2];
    bool qxy[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const int l = tid + 256 * e;
        const int qx = l % QX, qy = l / QX;
        qh[e] = qx + PX * qy;
        const int ex = ox + qx - 1, ey = oy + qy - 1;
        qxy[e] = l < QN && ex >= 0 && ex < nex && ey >= 0 && ey < ney;
        qo[e] = l;
    }
    const int hc = (tx + 1) + PX * (ty + 1);
    const int q0 = tx + QX * ty;
    lds_barrier();

    // ---- march through the planes ------------------------------------------
#pragma unroll
    for (int pl = 0; pl < NP; ++pl) {
        const int p = oz - 1 + pl;
        const bool pin = p >= 0 && p < nz;
        if (pl > 0) lds_barrier();
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int l = tid + 256 * e;
            if (l < PN) {
                float kv = 0.0f, cv = 0.0f;
                if (pin && eok[e]) {
                    if (KIN) {
                        kv = rK[pl][e];
                        cv = rS[pl][e];
                    } else {
                        const int idx = eoff[e] + nxy * p;
                        bool s1;
                        props_c(pc, sTab, rT[pl][e], rS[pl][e] > 0.499f, idx < a.n_sub, kv, cv,
                                s1);
                        if (a.S1out && p >= oz && l % PX >= 1 && l % PX <= LX && l / PX >= 1 &&
                            l / PX <= LY)
                            a.S1out[idx] = (uint8_t)s1;
                    }
                }
                sT[pl % 3][l] = rT[pl][e];
                sKn[l] = kv;
                sCn[l] = cv;
            }
        }
        lds_barrier();
        const bool ezok = p - 1 >= 0 && p - 1 < nez_act;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            if (qo[e] < QN) {
                const int h0 = qh[e], l = qo[e];
                const float Sk = (sKn[h0] + sKn[h0 + 1]) + (sKn[h0 + PX] + sKn[h0 + PX + 1]);
                const float Sc = (sCn[h0] + sCn[h0 + 1]) + (sCn[h0 + PX] + sCn[h0 + PX + 1]);
                const bool ok = qxy[e] && ezok;
                sKe[pl & 1][l] = ok ? (sSk[(pl + 1) & 1][l] + Sk) * 0.125f : 0.0f;
                sCe[pl & 1][l] = ok ? (sSc[(pl + 1) & 1][l] + Sc) * 0.125f : 0.0f;
                sSk[pl & 1][l] = Sk;
                sSc[pl & 1][l] = Sc;
            }
        }
        lds_barrier();
        if (pl < 2) continue;
        const int z = pl - 2;
        const int zo = oz + z;
        if (zo >= nz || !col_ok) continue;
        const int idx = cidx + nxy * zo;
        const float* Tm = sT[(pl + 1) % 3];
        const float* T0 = sT[(pl + 2) % 3];
        const float* Tp = sT[pl % 3];
        const float* kb = sKe[(pl + 1) & 1];  // elements (zo-1, zo)
        const float* ka = sKe[pl & 1];        // elements (zo, zo+1)
        const float* cb = sCe[(pl + 1) & 1];
        const float* ca = sCe[pl & 1];
        const float k000 = kb[q0], k100 = kb[q0 + 1], k010 = kb[q0 + QX], k110 = kb[q0 + QX + 1];
        const float k001 = ka[q0], k101 = ka[q0 + 1], k011 = ka[q0 + QX], k111 = ka[q0 + QX + 1];
        const float Msum = ((cb[q0] + cb[q0 + 1]) + (cb[q0 + QX] + cb[q0 + QX + 1])) +
                           ((ca[q0] + ca[q0 + 1]) + (ca[q0 + QX] + ca[q0 + QX + 1]));
        const float Pz00 = k000 + k001, Pz10 = k100 + k101, Pz01 = k010 + k011, Pz11 = k110 + k111;
        const float Py00 = k000 + k010, Py10 = k100 + k110, Py01 = k001 + k011, Py11 = k101 + k111;
        const float Px00 = k000 + k100, Px10 = k010 + k110, Px01 = k001 + k101, Px11 = k011 + k111;
        const float Fx0 = Pz00 + Pz01, Fx1 = Pz10 + Pz11;
        const float Fy0 = Pz00 + Pz10, Fy1 = Pz01 + Pz11;
        const float Fz0 = Py00 + Py10, Fz1 = Py01 + Py11;
        const float Kc = Fx0 + Fx1;
        const float Ti = T0[hc];
        float KT = a.kc[0] * (Kc * Ti);
        KT += a.kc[1] * (Fx0 * T0[hc - 1] + Fx1 * T0[hc + 1]);
        KT += a.kc[2] * (Fy0 * T0[hc - PX] + Fy1 * T0[hc + PX]);
        KT += a.kc[4] * (Fz0 * Tm[hc] + Fz1 * Tp[hc]);
        KT += a.kc[3] * (Pz00 * T0[hc - 1 - PX] + Pz10 * T0[hc + 1 - PX] + Pz01 * T0[hc - 1 + PX] +
                         Pz11 * T0[hc + 1 + PX]);
        KT += a.kc[5] * (Py00 * Tm[hc - 1] + Py10 * Tm[hc + 1] + Py01 * Tp[hc - 1] +
                         Py11 * Tp[hc + 1]);
        KT += a.kc[6] * (Px00 * Tm[hc - PX] + Px10 * Tm[hc + PX] + Px01 * Tp[hc - PX] +
                         Px11 * Tp[hc + PX]);
        KT += a.kc[7] * (k000 * Tm[hc - 1 - PX] + k100 * Tm[hc + 1 - PX] + k010 * Tm[hc - 1 + PX] +
                         k110 * Tm[hc + 1 + PX] + k001 * Tp[hc - 1 - PX] + k101 * Tp[hc + 1 - PX] +
                         k011 * Tp[hc - 1 + PX] + k111 * Tp[hc + 1 + PX]);
        const float M = Msum * mfac;
        float F = 0.0f;
        if (SRC) F = sSrc[tx] * sSrc[LX + ty] * sSrc[LX + LY + z] * a.w3;
        if (a.conv && zo == nez_act) {
            float cr = 0.0f;
            for (int e = 0; e < 4; ++e) {
                const int exx = ix - 1 + (e & 1), eyy = iy - 1 + (e >> 1);
                if (exx < 0 || exx >= nex || eyy < 0 || eyy >= ney) continue;
                const int f0 = (exx - ox + 1) + PX * (eyy - oy + 1);
                const float Tf[4] = {T0[f0], T0[f0 + 1], T0[f0 + 1 + PX], T0[f0 + PX]};
                const int ax = ix - exx, ay = iy - eyy;
                const int ai = ay ? (ax ? 2 : 3) : (ax ? 1 : 0);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float Tq = n2(q, 0) * Tf[0] + n2(q, 1) * Tf[1] + n2(q, 2) * Tf[2] +
                                     n2(q, 3) * Tf[3];
                    cr += n2(q, ai) * (face_flux(*m, Tq) * a.w2);
                }
            }
            F += cr;
        }
        float Tnew = ((M * Ti - KT) + F + rC[z]) / M;
        if (idx >= a.nn_active) Tnew = Tamb;
        if (BC == GM_BC_L1) {
            if (zo == 0) Tnew = a.bc_vals[4];
            else if (ix == nx - 1) Tnew = a.bc_vals[3];
            else if (ix == 0) Tnew = a.bc_vals[2];
            else if (iy == ny - 1) Tnew = a.bc_vals[1];
            else if (iy == 0) Tnew = a.bc_vals[0];
        } else if (BC == GM_BC_FINE) {
            if (face_col || zo == 0) Tnew = a.alpha * rFa[z] + (1.0f - a.alpha) * rFb[z];
        }
        if (a.clamp && Tnew < Tamb) Tnew = Tamb;
        a.Tout[idx] = Tnew;
        if (ACC) {
            const bool s2n = Ti >= pc.Tliq;
            const bool reset = (!rS2[z]) && s2n;
            const int lz0 = d->l0_lo3[2] + zo;
            if (l0xy && lz0 >= 0 && lz0 < d->l0_n[2]) {
                const int64_t l0 = (int64_t)l0base + (int64_t)d->l0_n[0] * d->l0_n[1] * lz0;
                const float racc = reset ? rAcc[z] : 0.0f;
                a.max_accum[l0] = racc > rMax[z] ? racc : rMax[z];
                a.accum[l0] = rAcc[z] + d->rows[a.src_row][5] * (s2n ? 1.0f : 0.0f) - racc;
            }
            a.S2out[idx] = (uint8_t)s2n;
        }
    }
}

template <bool S1F, bool SRC, int BC, bool ACC, bool KIN>
static void launch_ls(const LSArgs& a, const GmStepDesc* d, const gm_grid& hg, const DevMat* m,
                      hipStream_t s) {
    const int ntx = (hg.n[0] + LX - 1) / LX, nty = (hg.n[1] + LY - 1) / LY;
    // 4 planes per block on meshes up to ~1M nodes (enough blocks for 256
    // CUs), 8 planes on larger ones (less halo work per node)
    if ((int64_t)hg.n[0] * hg.n[1] * hg.n[2] > (1 << 20)) {
        const int nb = ntx * nty * ((hg.n[2] + 7) / 8);
        hipLaunchKernelGGL((k_level_step<S1F, SRC, BC, ACC, KIN, 8>), dim3(nb), dim3(256), 0, s, a,
                           d, m);
    } else {
        const int nb = ntx * nty * ((hg.n[2] + 3) / 4);
        hipLaunchKernelGGL((k_level_step<S1F, SRC, BC, ACC, KIN, 4>), dim3(nb), dim3(256), 0, s, a,
                           d, m);
    }
}

static void launch_step_l1(const LSArgs& a, const GmStepDesc* d, const gm_grid& hg,
                           const DevMat* m, hipStream_t s);

void launch_level_step(const LSArgs& a, const GmStepDesc* d, const gm_grid& hg, const DevMat* m,
                       hipStream_t s) {
    if (a.kin) {
        launch_ls<false, false, GM_BC_NONE, false, true>(a, d, hg, m, s);
    } else if (a.s1_float) {  // L1
        if (a.bc_mode == GM_BC_L1 && !a.S1out && !a.corr_full && !a.Fext &&
            a.src_mode == GM_SRC_NONE && (int64_t)hg.n[0] * hg.n[1] * hg.n[2] < ((int64_t)1 << 29)) {
            launch_step_l1(a, d, hg, m, s);
        } else if (a.bc_mode == GM_BC_L1) {
            launch_ls<true, false, GM_BC_L1, false, false>(a, d, hg, m, s);
        } else {
            launch_ls<true, false, GM_BC_NONE, false, false>(a, d, hg, m, s);
        }
    } else if (a.src_mode == GM_SRC_L3) {  // L3
        if (a.accum_on) launch_ls<false, true, GM_BC_FINE, true, false>(a, d, hg, m, s);
        else launch_ls<false, true, GM_BC_FINE, false, false>(a, d, hg, m, s);
    } else {  // L2
        launch_ls<false, false, GM_BC_FINE, false, false>(a, d, hg, m, s);
    }
}

// ---------------------------------------------------------------------------
// Fine-level (L2, L3) step of the subcycle, without LDS or barriers.
// The nodal properties are carried between steps: the input's k and cp come
// from the previous step (or from k_props_kc), so each node's table lookups
// happen once, where its temperature is produced:
//   S1 state   : S1 | (T >= T_liq) | substrate        (cF:2585-2592)
//   T+         : the explicit update of k_level_step   (cF:582-642), with the
//                time-blended coarse field on the Dirichlet faces
//   k+, cp+    : properties of T+ with S1 | (T+ >= T_liq), i.e. what the next
//                step's computeStateProperties returns.
// A wave owns DXS consecutive positions of the plane numbering
// q = x + nx * band (bands of YB rows), over ZC planes: lane l holds
// q = DXS s - 1 + l (lanes 0 and 63 only supply the halo).  Consecutive q
// run across band ends, so the lanes of a wave stay busy whatever nx is; the
// pair (nx-1, next band's 0) is no element, so its coefficients are zero.
// Every y and z neighbour of a lane's nodes is in its registers and x
// neighbours come from the adjacent lanes through DPP wave shifts.  Global
// memory goes through buffer resources with 32-bit offsets: positions
// outside the mesh read zero and their stores are dropped (offset past the
// end), so the body has no branches and the nodes of a brick interleave.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float lane_prev(float v) {  // value of lane l-1 (0 in lane 0)
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float lane_next(float v) {  // value of lane l+1 (0 in lane 63)
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));
}

#define DXS 62             // output positions per wave
#define GM_BUF_DW3 0x00020000  // buffer resource word 3 (gfx9 family, raw dword access)
#define GM_OOB 0x80000000u     // an offset past the end of every buffer

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, GM_BUF_DW3);
}
__device__ __forceinline__ float buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff = 0) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ void buf_st(float v, __amdgpu_buffer_rsrc_t r, uint32_t voff,
                                       uint32_t soff = 0) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, 0);
}
__device__ __forceinline__ uint32_t buf_ld8(__amdgpu_buffer_rsrc_t r, uint32_t voff) {
    return __builtin_amdgcn_raw_buffer_load_b8(r, voff, 0, 0);
}
__device__ __forceinline__ void buf_st8(uint32_t v, __amdgpu_buffer_rsrc_t r, uint32_t voff) {
    __builtin_amdgcn_raw_buffer_store_b8((uint8_t)v, r, voff, 0, 0);
}

// Top-face flux density of face_flux with hardware reciprocal, reciprocal
// square root and exponential (relative error a few 1e-7).
__device__ __forceinline__ float face_flux_hw(const DevMat& m, float Tq) {
    Tq = (Tq > m.Tclip) ? m.Tclip : Tq;
    float S = 0.0f;
    if (m.evc_cp != 0.0f) {  // evaporation (uniform: a zero coefficient skips it)
        const float invT = __builtin_amdgcn_rcpf(Tq);
        const float E = m.Lev + m.cp_f * (Tq - m.T_amb);
        const float mol = m.sqCM * __builtin_amdgcn_rsqf(Tq);
        S = m.evc_cp * __expf(-m.CT * (invT - m.invTb)) * mol * E;
    }
    const float t2 = Tq * Tq;
    const float q = m.h_conv * (m.T_amb - Tq) + m.sigma_eps * (m.Tamb4 - t2 * t2) - S;
    return q * 1e-6f;
}

// WPB waves per block, each its own brick (block b holds bricks
// WPB*tile(b) .. WPB*tile(b) + WPB - 1; bricks past the last are empty).
template <bool SRC, bool ACC, bool REC, int YB, int ZC, int WPB>
__global__ void __launch_bounds__(64 * WPB) k_step_dpp(LSArgs a, const GmStepDesc* __restrict__ d,
                                                       const DevMat* __restrict__ m) {
#define GM_STEP_BRICK \
    (WPB == 1 ? xcd_tile(blockIdx.x, gridDim.x) \
              : xcd_tile(blockIdx.x, gridDim.x) * WPB + (int)(threadIdx.x >> 6))
#define GM_STEP_LANE (WPB == 1 ? (int)threadIdx.x : (int)(threadIdx.x & 63))
#define GM_STEP_ROW d->rows[a.src_row < 0 ? 0 : a.src_row]
#include "gm_step_dpp_body.inc"
#undef GM_STEP_ROW
#undef GM_STEP_LANE
#undef GM_STEP_BRICK
}

// Item of block b in a batched launch: the last i < n with bstart[i] <= b
// (bstart non-decreasing from bstart[0] = 0, so an item with no blocks is
// skipped), and its first block *b0.  One load of 64 starts per lane group
// and a wave ballot instead of a chain of dependent loads: b is uniform, so
// every lane of the wave gets the same answer.  Called by every wave of the
// block.
__device__ __forceinline__ int batch_item_of(const int* __restrict__ bstart, int n, int b,
                                             int* b0) {
    const int lane = (int)(threadIdx.x & 63);
    int cnt = 0, vlast = 0, clast = 0;
    for (int c = 0; c < n; c += 64) {
        const int i = c + lane;
        const int v = i < n ? bstart[i] : 0x7fffffff;
        const unsigned long long bal = __ballot(v <= b);
        if (bal) {
            cnt += __popcll(bal);
            vlast = v;
            clast = c;
        }
    }
    const int lo = cnt - 1;
    *b0 = __builtin_amdgcn_readlane(vlast, lo - clast);
    return lo;
}

// Batched L3-type steps (source, surface flux, Dirichlet faces, no
// correction): many instances of their own mesh size and laser in one launch.
// bstart[i] is the first block of instance i (bstart[n] = total); mats[i] is
// its material (the same as it[i].m).  The blocks are dealt out XCD-aware as
// in the single launch, then each finds its instance (batch_item_of) and
// runs that instance's brick.  Items carry their time step by value
// (dt_mode GM_DT_VALUE); the desc's rows are read only under GM_DT_ROW.
// Both kernels include the body directly (no shared device function), so
// each compiles to the same instruction DAG as the single launch of the same
// template arguments and an instance's result equals its own launch bitwise.
#define GM_BATCH_FIND_INSTANCE                                           \
    const int b = xcd_tile(blockIdx.x, gridDim.x);                       \
    int b0_;                                                             \
    const int lo = batch_item_of(bstart, n, b, &b0_);                    \
    const LSArgs a = it[lo].a;                                           \
    /* materials indexed off the kernel argument (a known global pointer: \
       scalar loads), not through the pointer stored in the item */      \
    const DevMat* __restrict__ m = mats + lo;                            \
    const int tb = b - b0_;

template <bool REC, int YB, int ZC>
__global__ void __launch_bounds__(64) k_step_dpp_batch(const BatchItem* __restrict__ it,
                                                       const DevMat* __restrict__ mats,
                                                       const int* __restrict__ bstart, int n,
                                                       const GmStepDesc* __restrict__ d) {
    constexpr bool SRC = true, ACC = false;
    GM_BATCH_FIND_INSTANCE
#define GM_STEP_BRICK tb
#define GM_STEP_LANE (int)threadIdx.x
#define GM_STEP_ROW d->rows[a.src_row < 0 ? 0 : a.src_row]
#include "gm_step_dpp_body.inc"
#undef GM_STEP_ROW
#undef GM_STEP_LANE
#undef GM_STEP_BRICK
}

// The same with melt time (Part 2 of the subcycle), bricks of YB rows x ZC
// planes: the melt time adds the item's own toolpath row dt.
template <bool REC, int YB, int ZC>
__global__ void __launch_bounds__(64) k_step_dpp_batch_acc(const BatchItem* __restrict__ it,
                                                           const DevMat* __restrict__ mats,
                                                           const int* __restrict__ bstart, int n,
                                                           const GmStepDesc* __restrict__ d) {
    constexpr bool SRC = true, ACC = true;
    GM_BATCH_FIND_INSTANCE
#define GM_STEP_BRICK tb
#define GM_STEP_LANE (int)threadIdx.x
#define GM_STEP_ROW it[lo].row
#include "gm_step_dpp_body.inc"
#undef GM_STEP_ROW
#undef GM_STEP_LANE
#undef GM_STEP_BRICK
}
// The same for the L2 steps of the subcycle (no source, no melt time; the
// restricted terms from each item's FP64 node box), bricks of 1 row x 2
// planes as the single launch; items carry dt by value.
template <bool REC>
__global__ void __launch_bounds__(64) k_step_dpp_batch_l2(const BatchItem* __restrict__ it,
                                                          const DevMat* __restrict__ mats,
                                                          const int* __restrict__ bstart, int n,
                                                          const GmStepDesc* __restrict__ d) {
    constexpr bool SRC = false, ACC = false;
    constexpr int YB = 1, ZC = 2;
    GM_BATCH_FIND_INSTANCE
#define GM_STEP_BRICK tb
#define GM_STEP_LANE (int)threadIdx.x
#define GM_STEP_ROW d->rows[a.src_row < 0 ? 0 : a.src_row]
#include "gm_step_dpp_body.inc"
#undef GM_STEP_ROW
#undef GM_STEP_LANE
#undef GM_STEP_BRICK
}
#undef GM_BATCH_FIND_INSTANCE

// L1 step (computeL1Temperature, cF:2135-2204 with the L1 faces of
// cF:1568-1597) in the same wave-brick form: properties from (T, S1) of every
// loaded position (L1 carries no k, cp), active element layers below
// nez_act = ne_active / (nex ney), the surface flux on the top active plane,
// the restricted correction box, inactive nodes (idx >= nn_active) at T_amb
// and the Dirichlet values of the five L1 faces.
template <bool REC, int YB, int ZC>
__global__ void __launch_bounds__(64) k_step_l1(LSArgs a, const GmStepDesc* __restrict__ d,
                                                const DevMat* __restrict__ m) {
    const int nx = a.grid.n[0], ny = a.grid.n[1], nz = a.grid.n[2];
    const int nxy = nx * ny;
    const int nn = nxy * nz;
    const int nex = nx - 1, ney = ny - 1;
    const int nez_act = (int)(a.ne_active / ((int64_t)nex * ney));
    const int nband = (ny + YB - 1) / YB;
    const int nq = nx * nband;
    const int nzc = (nz + ZC - 1) / ZC;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int qs = fdiv_small(t, __builtin_amdgcn_rcpf((float)nzc)), zc = t - qs * nzc;
    const int lane = threadIdx.x;
    const int q = DXS * qs - 1 + lane;
    const bool qv = q >= 0 && q < nq;
    int band = 0;
    if (nq < (1 << 20)) band = qv ? fdiv_small(q, __builtin_amdgcn_rcpf((float)nx)) : 0;
    else band = qv ? (int)((unsigned)q / (unsigned)nx) : 0;
    const int x = qv ? q - band * nx : -1;
    const int y0 = band * YB, z0 = zc * ZC;
    const bool ein = qv && x < nex;
    const bool olane = lane >= 1 && lane <= DXS && qv;
    const int base = x + nx * y0 + nxy * z0;
    const uint32_t nb4 = (uint32_t)nn * 4u;
    const int nsub = (int)(a.n_sub < (int64_t)nn ? a.n_sub : (int64_t)nn);

    const auto rT = buf_rsrc(a.T, nb4), rS = buf_rsrc(a.S1, nb4);
    float Tw[ZC + 2][YB + 2], Sw[ZC + 2][YB + 2];
    bool rowok[YB + 2];
#pragma unroll
    for (int r = 0; r < YB + 2; ++r) {
        const int yy = y0 - 1 + r;
        rowok[r] = qv && yy >= 0 && yy < ny;
    }
#pragma unroll
    for (int p = 0; p < ZC + 2; ++p) {
        const int zz = z0 - 1 + p;
        const bool pok = zz >= 0 && zz < nz;
#pragma unroll
        for (int r = 0; r < YB + 2; ++r) {
            const uint32_t vo =
                (pok && rowok[r]) ? 4u * (uint32_t)(base + nx * (r - 1) + nxy * (p - 1)) : GM_OOB;
            Tw[p][r] = buf_ld(rT, vo);
            Sw[p][r] = buf_ld(rS, vo);
        }
    }
    float corr[ZC][YB];
#pragma unroll
    for (int i = 0; i < ZC; ++i)
#pragma unroll
        for (int j = 0; j < YB; ++j) corr[i][j] = 0.0f;
    if (a.corr) {
        const int c0x = a.corr_info[0][0], c1x = a.corr_info[0][1];
        const int c0y = a.corr_info[1][0], c1y = a.corr_info[1][1];
        const int c0z = a.corr_info[2][0], c1z = a.corr_info[2][1];
#pragma unroll
        for (int i = 0; i < ZC; ++i)
#pragma unroll
            for (int j = 0; j < YB; ++j) {
                const int bx = x - c0x, by = y0 + j - c0y, bz = z0 + i - c0z;
                if (olane && y0 + j < ny && z0 + i < nz && bx >= 0 && bx <= c1x && by >= 0 &&
                    by <= c1y && bz >= 0 && bz <= c1z)
                    corr[i][j] = a.corr[bx + a.corr_nmax[0] * (by + a.corr_nmax[1] * bz)];
            }
    }
    __shared__ float sTab[PTW] __attribute__((aligned(16)));
    stage_tables(m, sTab, lane, 64);
    float dt = a.dt;
    if (a.dt_mode == GM_DT_ROW) dt = d->rows[a.dt_index][5];
    else if (a.dt_mode == GM_DT_L1) dt = d->dt1;
    else if (a.dt_mode == GM_DT_L2) dt = d->dt2[a.dt_index];
    const float rmfac = dt * __builtin_amdgcn_rcpf(a.mrow);  // uniform: one reciprocal
    const PropC pc = load_pc(m);
    const float Tamb = m->T_amb;

    lds_barrier();  // property table staged
    // properties of every loaded position (computeStateProperties, cF:2567-2614)
    float Kw[ZC + 2][YB + 2], Cw[ZC + 2][YB + 2];
#pragma unroll
    for (int p = 0; p < ZC + 2; ++p) {
#pragma unroll
        for (int r = 0; r < YB + 2; ++r) {
            const int idx = base + nx * (r - 1) + nxy * (p - 1);
            bool s1;
            if (REC) props_rec(pc, sTab, Tw[p][r], Sw[p][r] > 0.499f, idx < nsub, Kw[p][r], Cw[p][r], s1);
            else props_c(pc, sTab, Tw[p][r], Sw[p][r] > 0.499f, idx < nsub, Kw[p][r], Cw[p][r], s1);
        }
    }
    // element means over the active layers
    float Qk[ZC + 2][YB + 1], Qc[ZC + 2][YB + 1];
#pragma unroll
    for (int p = 0; p < ZC + 2; ++p) {
#pragma unroll
        for (int r = 0; r < YB + 1; ++r) {
            const int yy = y0 - 1 + r;
            const bool ok = ein && yy >= 0 && yy < ney;
            const float ak = Kw[p][r] + Kw[p][r + 1];
            const float ac = Cw[p][r] + Cw[p][r + 1];
            const float qk = ak + lane_next(ak);
            const float qc = ac + lane_next(ac);
            Qk[p][r] = ok ? qk : 0.0f;
            Qc[p][r] = ok ? qc : 0.0f;
        }
    }
    float Ek[ZC + 1][YB + 1], Ec[ZC + 1][YB + 1];
#pragma unroll
    for (int p = 0; p < ZC + 1; ++p) {
        const int ez = z0 - 1 + p;
        const float f = (ez >= 0 && ez < nez_act) ? 0.125f : 0.0f;
#pragma unroll
        for (int r = 0; r < YB + 1; ++r) {
            Ek[p][r] = (Qk[p][r] + Qk[p + 1][r]) * f;
            Ec[p][r] = (Qc[p][r] + Qc[p + 1][r]) * f;
        }
    }
    // surface flux of the top active plane (cF:2207-2301)
    float Ftop[YB];
#pragma unroll
    for (int j = 0; j < YB; ++j) Ftop[j] = 0.0f;
    const int ptop = nez_act - (z0 - 1);
    if (a.conv && ptop >= 1 && ptop <= ZC) {
        float Fq[YB + 1][4];
#pragma unroll
        for (int r = 0; r < YB + 1; ++r) {
            float t0 = 0.0f, t3 = 0.0f;
#pragma unroll
            for (int p = 1; p <= ZC; ++p)
                if (p == ptop) {
                    t0 = Tw[p][r];
                    t3 = Tw[p][r + 1];
                }
            const float t1 = lane_next(t0), t2 = lane_next(t3);
            const int yy = y0 - 1 + r;
            const bool ok = ein && yy >= 0 && yy < ney;
#pragma unroll
            for (int qd = 0; qd < 4; ++qd) {
                const float Tq = n2(qd, 0) * t0 + n2(qd, 1) * t1 + n2(qd, 2) * t2 + n2(qd, 3) * t3;
                const float fl = face_flux_hw(*m, Tq) * a.w2;
                Fq[r][qd] = ok ? fl : 0.0f;
            }
        }
#pragma unroll
        for (int j = 0; j < YB; ++j) {
            float cr = 0.0f;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int bxe = e & 1, bye = e >> 1;
                const int ai = bye ? (bxe ? 0 : 1) : (bxe ? 3 : 2);
#pragma unroll
                for (int qd = 0; qd < 4; ++qd) {
                    const float v = Fq[j + bye][qd];
                    cr += n2(qd, ai) * (bxe ? v : lane_prev(v));
                }
            }
            Ftop[j] = cr;
        }
    }

    const auto rTo = buf_rsrc(a.Tout, nb4);
#pragma unroll
    for (int i = 0; i < ZC; ++i) {
#pragma unroll
        for (int j = 0; j < YB; ++j) {
            const int yy = y0 + j, zz = z0 + i;
            const float k100 = Ek[i][j], k110 = Ek[i][j + 1];
            const float k101 = Ek[i + 1][j], k111 = Ek[i + 1][j + 1];
            const float k000 = lane_prev(k100), k010 = lane_prev(k110);
            const float k001 = lane_prev(k101), k011 = lane_prev(k111);
            const float cs = (Ec[i][j] + Ec[i][j + 1]) + (Ec[i + 1][j] + Ec[i + 1][j + 1]);
            const float Msum = cs + lane_prev(cs);
            const float Ti = Tw[i + 1][j + 1];
            const float Pz00 = k000 + k001, Pz10 = k100 + k101, Pz01 = k010 + k011,
                        Pz11 = k110 + k111;
            const float Py00 = k000 + k010, Py10 = k100 + k110, Py01 = k001 + k011,
                        Py11 = k101 + k111;
            const float Px00 = k000 + k100, Px10 = k010 + k110, Px01 = k001 + k101,
                        Px11 = k011 + k111;
            const float Fx0 = Pz00 + Pz01, Fx1 = Pz10 + Pz11;
            const float Fy0 = Pz00 + Pz10, Fy1 = Pz01 + Pz11;
            const float Fz0 = Py00 + Py10, Fz1 = Py01 + Py11;
            float KT = a.kc[0] * ((Fx0 + Fx1) * Ti);
            KT += a.kc[1] * (Fx0 * lane_prev(Ti) + Fx1 * lane_next(Ti));
            KT += a.kc[2] * (Fy0 * Tw[i + 1][j] + Fy1 * Tw[i + 1][j + 2]);
            KT += a.kc[4] * (Fz0 * Tw[i][j + 1] + Fz1 * Tw[i + 2][j + 1]);
            KT += a.kc[3] * (Pz00 * lane_prev(Tw[i + 1][j]) + Pz10 * lane_next(Tw[i + 1][j]) +
                             Pz01 * lane_prev(Tw[i + 1][j + 2]) + Pz11 * lane_next(Tw[i + 1][j + 2]));
            KT += a.kc[5] * (Py00 * lane_prev(Tw[i][j + 1]) + Py10 * lane_next(Tw[i][j + 1]) +
                             Py01 * lane_prev(Tw[i + 2][j + 1]) + Py11 * lane_next(Tw[i + 2][j + 1]));
            KT += a.kc[6] * (Px00 * Tw[i][j] + Px10 * Tw[i][j + 2] + Px01 * Tw[i + 2][j] +
                             Px11 * Tw[i + 2][j + 2]);
            KT += a.kc[7] * (k000 * lane_prev(Tw[i][j]) + k100 * lane_next(Tw[i][j]) +
                             k010 * lane_prev(Tw[i][j + 2]) + k110 * lane_next(Tw[i][j + 2]) +
                             k001 * lane_prev(Tw[i + 2][j]) + k101 * lane_next(Tw[i + 2][j]) +
                             k011 * lane_prev(Tw[i + 2][j + 2]) + k111 * lane_next(Tw[i + 2][j + 2]));
            float F = corr[i][j];
            if (i + 1 == ptop) F += Ftop[j];
            float Tnew = Ti + (F - KT) * (__builtin_amdgcn_rcpf(Msum) * rmfac);
            const int idx = base + nx * j + nxy * i;
            if (idx >= a.nn_active) Tnew = Tamb;
            if (zz == 0) Tnew = a.bc_vals[4];
            else if (x == nex) Tnew = a.bc_vals[3];
            else if (x == 0) Tnew = a.bc_vals[2];
            else if (yy == ney) Tnew = a.bc_vals[1];
            else if (yy == 0) Tnew = a.bc_vals[0];
            if (a.clamp && Tnew < Tamb) Tnew = Tamb;
            const bool ok = olane && yy < ny && zz < nz;
            buf_st(Tnew, rTo, ok ? 4u * (uint32_t)idx : GM_OOB);
        }
    }
}

static void launch_step_l1(const LSArgs& a0, const GmStepDesc* d, const gm_grid& hg,
                           const DevMat* m, hipStream_t s) {
    constexpr int YB = 1, ZC = 3;
    LSArgs a = a0;
    a.grid = hg;  // node counts (fixed for a level)
    const int nq = hg.n[0] * ((hg.n[1] + YB - 1) / YB);
    const int nb = ((nq + DXS - 1) / DXS) * ((hg.n[2] + ZC - 1) / ZC);
    if (a.prec_ok)
        hipLaunchKernelGGL((k_step_l1<true, YB, ZC>), dim3(nb), dim3(64), 0, s, a, d, m);
    else
        hipLaunchKernelGGL((k_step_l1<false, YB, ZC>), dim3(nb), dim3(64), 0, s, a, d, m);
}

// waves (bricks) per block of the subcycle steps.  Measured on the
// production deck on two machines: four per block ran the L3 step in
// 6.69 us on both, one per block in 6.87 us on one and 6.51 us on the other,
// where it also made the whole L1 step 6% faster (2.14 ms against 2.27); one
// per block is kept for all levels.
#ifndef GM_L3_WPB
#define GM_L3_WPB 1
#endif
#ifndef GM_L3ACC_WPB
#define GM_L3ACC_WPB GM_L3_WPB
#endif
#ifndef GM_L2_WPB
#define GM_L2_WPB 1
#endif

template <bool SRC, bool ACC, bool REC>
static void launch_dpp(const LSArgs& a, const GmStepDesc* d, const gm_grid& hg, const DevMat* m,
                       hipStream_t s) {
    // measured best on 101 x 101 x 51 meshes: 3 planes per brick for L3
    // steps, 2 for L2 steps (whose correction-box reads favour more waves)
    constexpr int YB = 1, ZC = SRC ? 3 : 2;
    constexpr int WPB = SRC ? (ACC ? GM_L3ACC_WPB : GM_L3_WPB) : GM_L2_WPB;
    const int nq = hg.n[0] * ((hg.n[1] + YB - 1) / YB);
    const int nbr = ((nq + DXS - 1) / DXS) * ((hg.n[2] + ZC - 1) / ZC);
    const int nb = (nbr + WPB - 1) / WPB;
    hipLaunchKernelGGL((k_step_dpp<SRC, ACC, REC, YB, ZC, WPB>), dim3(nb), dim3(64 * WPB), 0, s,
                       a, d, m);
}

void launch_level_step_kc(const LSArgs& a0, const GmStepDesc* d, const gm_grid& hg,
                          const DevMat* m, hipStream_t s) {
    LSArgs a = a0;
    a.grid = hg;  // the kernel takes the node counts from here
    if (!a.kcin || !a.kcout) return;  // no carried properties for these buffers
    if (a.src_mode == GM_SRC_L3 && !a.srcT) return;  // the source comes from tables
    // the buffer offsets are 32-bit: node counts up to 2^28
    if ((int64_t)hg.n[0] * hg.n[1] * hg.n[2] * 8 >= (int64_t)1 << 31) return;
    if (a.src_mode == GM_SRC_L3) {
        if (a.accum_on) {
            if (a.prec_ok) launch_dpp<true, true, true>(a, d, hg, m, s);
            else launch_dpp<true, true, false>(a, d, hg, m, s);
        } else {
            if (a.prec_ok) launch_dpp<true, false, true>(a, d, hg, m, s);
            else launch_dpp<true, false, false>(a, d, hg, m, s);
        }
    } else {
        if (a.prec_ok) launch_dpp<false, false, true>(a, d, hg, m, s);
        else launch_dpp<false, false, false>(a, d, hg, m, s);
    }
}

// ---- batched L3-type steps -------------------------------------------------
int batch_bricks(const gm_grid& g, int yb, int zc) {
    const int nq = g.n[0] * ((g.n[1] + yb - 1) / yb);
    return ((nq + DXS - 1) / DXS) * ((g.n[2] + zc - 1) / zc);
}

// brick shapes (rows per band x planes) the batched step offers its autotuner
const int kBatchTiles[GM_BATCH_TILES][2] = {{1, 2}, {1, 3}, {1, 4}, {2, 2}, {2, 3}};

// separable source factors of each instance's row on its own mesh (as
// k_src3_tables: x factors times 6 sqrt(3) eta P w3, then y, z); blockIdx.y
// is the instance
__global__ void k_batch_src(const BatchItem* __restrict__ it) {
    const BatchItem& b = it[blockIdx.y];
    const gm_grid& g = b.a.grid;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.n[0] + g.n[1] + g.n[2]) return;
    const DevMat* m = b.m;
    float v;
    if (i < g.n[0]) v = src_axis(g, 0, i, b.row[0], m->rsq, m->rc) * (m->pc6 * b.row[6] * b.w3);
    else if (i < g.n[0] + g.n[1]) v = src_axis(g, 1, i - g.n[0], b.row[1], m->rsq, m->rc);
    else v = src_axis(g, 2, i - g.n[0] - g.n[1], b.row[2], m->dsq, m->dc);
    ((float*)b.a.srcT)[i] = v;
}

void launch_batch_src(const BatchItem* it, int n, int maxaxes, hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(k_batch_src, dim3((maxaxes + 255) / 256, n), dim3(256), 0, s, it);
}

template <int YB, int ZC>
static void batch_step_t(const BatchItem* it, const DevMat* mats, const int* bstart, int n,
                         int nblocks, bool rec, const GmStepDesc* d, hipStream_t s) {
    if (rec)
        hipLaunchKernelGGL((k_step_dpp_batch<true, YB, ZC>), dim3(nblocks), dim3(64), 0, s, it,
                           mats, bstart, n, d);
    else
        hipLaunchKernelGGL((k_step_dpp_batch<false, YB, ZC>), dim3(nblocks), dim3(64), 0, s, it,
                           mats, bstart, n, d);
}

template <int YB, int ZC>
static void batch_step_acc_t(const BatchItem* it, const DevMat* mats, const int* bstart, int n,
                             int nblocks, bool rec, const GmStepDesc* d, hipStream_t s) {
    if (rec)
        hipLaunchKernelGGL((k_step_dpp_batch_acc<true, YB, ZC>), dim3(nblocks), dim3(64), 0, s, it,
                           mats, bstart, n, d);
    else
        hipLaunchKernelGGL((k_step_dpp_batch_acc<false, YB, ZC>), dim3(nblocks), dim3(64), 0, s,
                           it, mats, bstart, n, d);
}

void launch_batch_step_acc(const BatchItem* it, const DevMat* mats, const int* bstart, int n,
                           int nblocks, int tile, bool rec, const GmStepDesc* d, hipStream_t s) {
    if (n <= 0 || nblocks <= 0) return;
    switch (tile) {
        case 0: batch_step_acc_t<1, 2>(it, mats, bstart, n, nblocks, rec, d, s); break;
        case 2: batch_step_acc_t<1, 4>(it, mats, bstart, n, nblocks, rec, d, s); break;
        case 3: batch_step_acc_t<2, 2>(it, mats, bstart, n, nblocks, rec, d, s); break;
        case 4: batch_step_acc_t<2, 3>(it, mats, bstart, n, nblocks, rec, d, s); break;
        default: batch_step_acc_t<1, 3>(it, mats, bstart, n, nblocks, rec, d, s); break;
    }
}

void launch_batch_step_l2(const BatchItem* it, const DevMat* mats, const int* bstart, int n,
                          int nblocks, bool rec, const GmStepDesc* d, hipStream_t s) {
    if (n <= 0 || nblocks <= 0) return;
    if (rec)
        hipLaunchKernelGGL(k_step_dpp_batch_l2<true>, dim3(nblocks), dim3(64), 0, s, it, mats,
                           bstart, n, d);
    else
        hipLaunchKernelGGL(k_step_dpp_batch_l2<false>, dim3(nblocks), dim3(64), 0, s, it, mats,
                           bstart, n, d);
}

void launch_batch_step(const BatchItem* it, const DevMat* mats, const int* bstart, int n,
                       int nblocks, int tile, bool rec, const GmStepDesc* d, hipStream_t s) {
    if (n <= 0 || nblocks <= 0) return;
    switch (tile) {
        case 0: batch_step_t<1, 2>(it, mats, bstart, n, nblocks, rec, d, s); break;
        case 2: batch_step_t<1, 4>(it, mats, bstart, n, nblocks, rec, d, s); break;
        case 3: batch_step_t<2, 2>(it, mats, bstart, n, nblocks, rec, d, s); break;
        case 4: batch_step_t<2, 3>(it, mats, bstart, n, nblocks, rec, d, s); break;
        default: batch_step_t<1, 3>(it, mats, bstart, n, nblocks, rec, d, s); break;
    }
}

// the same step of one instance in a launch of its own (the unbatched form)
template <int YB, int ZC>
static void step_t(const LSArgs& a, const GmStepDesc* d, const gm_grid& hg, const DevMat* m,
                   bool rec, hipStream_t s) {
    const int nb = batch_bricks(hg, YB, ZC);
    if (rec)
        hipLaunchKernelGGL((k_step_dpp<true, false, true, YB, ZC, 1>), dim3(nb), dim3(64), 0, s, a, d,
                           m);
    else
        hipLaunchKernelGGL((k_step_dpp<true, false, false, YB, ZC, 1>), dim3(nb), dim3(64), 0, s, a,
                           d, m);
}

void launch_step_tile(const LSArgs& a0, const GmStepDesc* d, const gm_grid& hg, const DevMat* m,
                      int tile, bool rec, hipStream_t s) {
    LSArgs a = a0;
    a.grid = hg;
    switch (tile) {
        case 0: step_t<1, 2>(a, d, hg, m, rec, s); break;
        case 2: step_t<1, 4>(a, d, hg, m, rec, s); break;
        case 3: step_t<2, 2>(a, d, hg, m, rec, s); break;
        case 4: step_t<2, 3>(a, d, hg, m, rec, s); break;
        default: step_t<1, 3>(a, d, hg, m, rec, s); break;
    }
}

// k, cp of (T, S1) as the next level step's computeStateProperties returns them;
// blockIdx.y picks one of up to two fields
struct PropsJob {
    const float* T;
    const uint8_t* S1;
    int64_t nn, n_sub;
    float* kc;
};
struct PropsJobs {
    PropsJob j[2];
};

__global__ void k_props_kc(const DevMat* __restrict__ m, PropsJobs jobs) {
    __shared__ float sTab[PTW] __attribute__((aligned(16)));
    stage_tables(m, sTab, threadIdx.x, blockDim.x);
    __syncthreads();
    const PropsJob& J = jobs.j[blockIdx.y];
    PropC pc = load_pc(m);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < J.nn;
         i += (int64_t)gridDim.x * blockDim.x) {
        float kk, cc;
        bool s1;
        props_c(pc, sTab, J.T[i], J.S1[i] != 0, i < J.n_sub, kk, cc, s1);
        J.kc[i] = kk;
        J.kc[J.nn + i] = cc;
    }
}

void launch_props_kc(const DevMat* m, const float* T, const uint8_t* S1, int64_t nn,
                     int64_t n_sub, float* kc, hipStream_t s, const float* T2,
                     const uint8_t* S12, int64_t nn2, int64_t n_sub2, float* kc2) {
    PropsJobs jobs;
    jobs.j[0] = PropsJob{T, S1, nn, n_sub, kc};
    jobs.j[1] = PropsJob{T2, S12, nn2, n_sub2, kc2};
    const int nj = T2 ? 2 : 1;
    const int64_t nb = std::min<int64_t>((std::max(nn, T2 ? nn2 : 0) + 255) / 256, 4096);
    hipLaunchKernelGGL(k_props_kc, dim3((unsigned)nb, nj), dim3(256), 0, s, m, jobs);
}

// Up to 8 device-to-device copies in one launch (blockIdx.y picks the copy):
// 16-byte chunks, then the tail bytes.
__global__ void k_copy_multi(CopyList L) {
    const CopyItem& c = L.c[blockIdx.y];
    if ((((uintptr_t)c.dst) | ((uintptr_t)c.src)) & 15) {  // not 16-byte aligned: bytes
        for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < c.bytes;
             i += (int64_t)gridDim.x * blockDim.x)
            ((uint8_t*)c.dst)[i] = ((const uint8_t*)c.src)[i];
        return;
    }
    const int64_t n16 = c.bytes >> 4;
    const uint4* src = (const uint4*)c.src;
    uint4* dst = (uint4*)c.dst;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n16;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
    const int64_t tail = c.bytes - (n16 << 4);
    if (blockIdx.x == 0 && threadIdx.x < tail)
        ((uint8_t*)c.dst)[(n16 << 4) + threadIdx.x] = ((const uint8_t*)c.src)[(n16 << 4) + threadIdx.x];
}

void launch_copy_multi(const CopyList& L, int n, hipStream_t s) {
    int64_t mx = 0;
    for (int i = 0; i < n; ++i) mx = std::max(mx, L.c[i].bytes >> 4);
    const int64_t nb = std::max<int64_t>(1, std::min<int64_t>((mx + 255) / 256, 2048));
    if (n > 0) hipLaunchKernelGGL(k_copy_multi, dim3((unsigned)nb, n), dim3(256), 0, s, L);
}

// Separable source factors of every L3 node line for the desc rows:
// src_axis of x (times 6 sqrt(3) eta P w3), y and z (cF:1013-1025, 2986-3010).
__global__ void k_src3_tables(const GmStepDesc* __restrict__ d, const DevMat* __restrict__ m,
                              int nrows, float w3, float* __restrict__ out) {
    const gm_grid& g = d->g[3];
    const int n = g.n[0] + g.n[1] + g.n[2];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrows * n) return;
    const int r = t / n, i = t % n;
    const float* row = d->rows[r];
    float v;
    if (i < g.n[0]) {
        v = src_axis(g, 0, i, row[0], m->rsq, m->rc) * (m->pc6 * row[6] * w3);
    } else if (i < g.n[0] + g.n[1]) {
        v = src_axis(g, 1, i - g.n[0], row[1], m->rsq, m->rc);
    } else {
        v = src_axis(g, 2, i - g.n[0] - g.n[1], row[2], m->dsq, m->dc);
    }
    out[t] = v;
}

void launch_src3_tables(const GmStepDesc* d, const DevMat* m, int nrows, int ntot, float w3,
                        float* out, hipStream_t s) {
    const int n = nrows * ntot;
    if (n > 0)
        hipLaunchKernelGGL(k_src3_tables, dim3((n + 255) / 256), dim3(256), 0, s, d, m, nrows, w3,
                           out);
}

// Melt-time state between Level0 and the L3 nodes (the L3 window maps one to
// one onto Level0 nodes, cF:3568-3578): to_l3 = 1 gathers, 0 scatters back
// (nodes of the window outside Level0 are left alone / read 0).
__global__ void k_acc_l3(const GmStepDesc* __restrict__ d, float* __restrict__ accum,
                         float* __restrict__ max_accum, float* __restrict__ acc3,
                         float* __restrict__ macc3, int64_t n3, int to_l3) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n3) return;
    const gm_grid& g3 = d->g[3];
    int ix, iy, iz;
    dec3(t, g3.n[0], g3.n[1], ix, iy, iz);
    const int lx = d->l0_lo3[0] + ix, ly = d->l0_lo3[1] + iy, lz = d->l0_lo3[2] + iz;
    const bool in0 = lx >= 0 && lx < d->l0_n[0] && ly >= 0 && ly < d->l0_n[1] && lz >= 0 &&
                     lz < d->l0_n[2];
    const int64_t l0 = in0 ? nid(lx, ly, lz, d->l0_n[0], d->l0_n[1]) : 0;
    if (to_l3) {
        acc3[t] = in0 ? accum[l0] : 0.0f;
        macc3[t] = in0 ? max_accum[l0] : 0.0f;
    } else if (in0) {
        accum[l0] = acc3[t];
        max_accum[l0] = macc3[t];
    }
}

void launch_acc_l3(const GmStepDesc* d, float* accum, float* max_accum, float* acc3, float* macc3,
                   int64_t n3, int to_l3, hipStream_t s) {
    hipLaunchKernelGGL(k_acc_l3, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, s, d, accum,
                       max_accum, acc3, macc3, n3, to_l3);
}

// Dirichlet values on the W, E, S, N and bottom faces of a fine level:
// A and B prolonged separately (interpolatePointsMatrix rule, cF:1060-1104)
// onto full-size arrays at the face nodes (cF:1598-1620); the level step
// blends them as alpha*A + (1-alpha)*B.
__device__ __forceinline__ void face_bc_at(const GmStepDesc* __restrict__ d, int fine,
                                           int coarse, const float* __restrict__ A,
                                           const float* __restrict__ B, float* __restrict__ fa,
                                           float* __restrict__ fb, int t) {
    const gm_grid& g = d->g[fine];
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const int nW = ny * nz, nS = nx * nz, nB = nx * ny;
    int ix, iy, iz;
    if (t < 2 * nW) {
        ix = t < nW ? 0 : nx - 1;
        t %= nW;
        iy = t % ny;
        iz = t / ny;
    } else if ((t -= 2 * nW) < 2 * nS) {
        iy = t < nS ? 0 : ny - 1;
        t %= nS;
        ix = t % nx;
        iz = t / nx;
    } else if ((t -= 2 * nS) < nB) {
        iz = 0;
        ix = t % nx;
        iy = t / nx;
    } else {
        return;
    }
    float N[8];
    int64_t nd[8];
    interp8(d->g[coarse], coordf(g, 0, ix), coordf(g, 1, iy), coordf(g, 2, iz), N, nd);
    float sa = 0.0f, sb = 0.0f;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        sa += N[q] * A[nd[q]];
        sb += N[q] * B[nd[q]];
    }
    const int idx = ix + nx * iy + nx * ny * iz;
    fa[idx] = sa;
    fb[idx] = sb;
}

__global__ void k_face_bc(const GmStepDesc* __restrict__ d, int fine, int coarse,
                          const float* __restrict__ A, const float* __restrict__ B,
                          float* __restrict__ fa, float* __restrict__ fb) {
    face_bc_at(d, fine, coarse, A, B, fa, fb, blockIdx.x * blockDim.x + threadIdx.x);
}

// Face values of many engines' L3 windows from their L2 fields in one
// launch: bstart[i] is the first block of item i (256 threads per block).
__global__ void __launch_bounds__(256) k_face_bc_batch(const FaceItem* __restrict__ it,
                                                       const int* __restrict__ bstart, int n) {
    const int b = (int)blockIdx.x;
    int b0;
    const int lo = batch_item_of(bstart, n, b, &b0);
    const FaceItem x = it[lo];
    face_bc_at(x.d, 3, 2, x.A, x.B, x.fa, x.fb, (b - b0) * 256 + (int)threadIdx.x);
}

void launch_face_bc_batch(const FaceItem* it, const int* bstart, int n, int nblocks,
                          hipStream_t s) {
    if (n <= 0 || nblocks <= 0) return;
    hipLaunchKernelGGL(k_face_bc_batch, dim3(nblocks), dim3(256), 0, s, it, bstart, n);
}

int face_bc_blocks(const gm_grid& gf) {
    const int n = 2 * gf.n[1] * gf.n[2] + 2 * gf.n[0] * gf.n[2] + gf.n[0] * gf.n[1];
    return (n + 255) / 256;
}

void launch_face_bc(const GmStepDesc* d, int fine, int coarse, const gm_grid& gf, const float* A,
                    const float* B, float* fa, float* fb, hipStream_t s) {
    const int n = 2 * gf.n[1] * gf.n[2] + 2 * gf.n[0] * gf.n[2] + gf.n[0] * gf.n[1];
    hipLaunchKernelGGL(k_face_bc, dim3((n + 255) / 256), dim3(256), 0, s, d, fine, coarse, A, B,
                       fa, fb);
}


// ---------------------------------------------------------------------------
// Fine -> coarse axis tables (the separable factors of the Shapes tensors)
// ---------------------------------------------------------------------------

__device__ __forceinline__ int coarse_elem(const gm_grid& gc, int d, float xq) {
    int c = (int)floorf((xq - gc.x0[d]) / gc.h[d]);
    const int nec = gc.n[d] - 1;
    return c < 0 ? 0 : (c > nec - 1 ? nec - 1 : c);
}

__device__ __forceinline__ float gauss_x(const gm_grid& gf, int d, int e, int gg) {
    const float xa = coordf(gf, d, e), xb = coordf(gf, d, e + 1);
    return gg == 0 ? (xa * GA + xb * GB) : (xa * GB + xb * GA);
}

struct GmPairs3 {
    GmPair p[3];
};

__global__ void k_axis_tables(GmPairs3 ps, const GmStepDesc* __restrict__ d) {
    const GmPair& p = ps.p[blockIdx.y];
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int ax = 0;
    while (ax < 3 && t >= p.nef[ax]) {
        t -= p.nef[ax];
        ++ax;
    }
    if (ax >= 3) return;
    const gm_grid& gf = d->g[p.fine];
    const gm_grid& gc = d->g[p.coarse];
    const int nef = p.nef[ax];
    const GmAxisTab& tb = p.ax[ax];
    float4 qq[2];
    for (int gg = 0; gg < 2; ++gg) {
        const float xq = gauss_x(gf, ax, t, gg);
        const int c = coarse_elem(gc, ax, xq);
        const float x0 = coordf(gc, ax, c), x1 = coordf(gc, ax, c + 1);
        qq[gg] = make_float4(xq, (x1 - xq) / gc.h[ax], (xq - x0) / gc.h[ax], 0.0f);
    }
    tb.q[2 * t] = qq[0];
    tb.q[2 * t + 1] = qq[1];
    const int c = coarse_elem(gc, ax, gauss_x(gf, ax, t, 0));
    tb.c0[t] = c;
    const int clo = coarse_elem(gc, ax, gauss_x(gf, ax, 0, 0));
    const int chi = coarse_elem(gc, ax, gauss_x(gf, ax, nef - 1, 0));
    const int k = c - clo;
    const int cmax = p.cmax[ax];
    if (k >= 0 && k < cmax) {
        if (t == 0 || coarse_elem(gc, ax, gauss_x(gf, ax, t - 1, 0)) != c) tb.info[2 + k] = t;
        if (t == nef - 1 || coarse_elem(gc, ax, gauss_x(gf, ax, t + 1, 0)) != c)
            tb.info[2 + cmax + k] = t + 1;
    }
    if (t == 0) {
        tb.info[0] = clo;
        const int n = chi - clo + 1;
        tb.info[1] = n < cmax ? n : cmax;
    }
}

void launch_axis_tables(const GmPair* p, int np, const GmStepDesc* d, hipStream_t s) {
    GmPairs3 ps;
    int n = 0;
    for (int i = 0; i < np && i < 3; ++i) {
        ps.p[i] = p[i];
        n = std::max(n, p[i].nef[0] + p[i].nef[1] + p[i].nef[2]);
    }
    hipLaunchKernelGGL(k_axis_tables, dim3((n + 255) / 256, np), dim3(256), 0, s, ps, d);
}

// ---------------------------------------------------------------------------
// Coarse source tables (computeLevelSource, cF:2667-2730, separably):
//   F_A = (w_f / npos) sum_pos Gx[pos][Ax] Gy[pos][Ay] Gz[pos][Az]
//   Gd[pos][A] = sum over fine elements e with c0(e) + a = A, Gauss g:
//                phi(e, g, a) * coef_d exp(-3 (xq - v_d)^2 / sq_d)
// (the x table also carries 6 sqrt(3) P eta).
// ---------------------------------------------------------------------------

__global__ void k_source_tables(GmPair p, const GmStepDesc* __restrict__ d, DevMat m, int row0,
                                int npos, int use_rows, const float* __restrict__ rows_dev,
                                float* __restrict__ G, int nmax) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npos * 3 * nmax) return;
    const int A = t % nmax, ax = (t / nmax) % 3, pos = t / (3 * nmax);
    const GmAxisTab& tb = p.ax[ax];
    const int cmax = p.cmax[ax];
    const int cn = tb.info[1];
    float s = 0.0f;
    if (A <= cn) {
        const float* row = use_rows ? rows_dev + 7 * (row0 + pos) : d->rows[row0 + pos];
        const float v = row[ax];
        const float sq = ax == 2 ? m.dsq : m.rsq;
        const float coef = ax == 2 ? m.dc : m.rc;
        for (int side = 0; side < 2; ++side) {
            const int c = A - 1 + side;  // element below (local node 1) / above (node 0)
            if (c < 0 || c >= cn) continue;
            const int e0 = tb.info[2 + c], e1 = tb.info[2 + cmax + c];
            for (int e = e0; e < e1; ++e) {
#pragma unroll
                for (int gg = 0; gg < 2; ++gg) {
                    const float4 q = tb.q[2 * e + gg];
                    const float tt = q.x - v;
                    const float qv = coef * expf((-3.0f * (tt * tt)) / sq);
                    s += (side == 0 ? q.z : q.y) * qv;
                }
            }
        }
        if (ax == 0) s *= m.pc6 * row[6];
    }
    G[t] = s;
}

void launch_source_tables(const GmPair& p, const GmStepDesc* d, const DevMat& m, int row0,
                          int npos, int use_rows, const float* rows_dev, float* G, int nmax,
                          hipStream_t s) {
    const int n = npos * 3 * nmax;
    hipLaunchKernelGGL(k_source_tables, dim3((n + 255) / 256), dim3(256), 0, s, p, d, m, row0,
                       npos, use_rows, rows_dev, G, nmax);
}

// ---------------------------------------------------------------------------
// Restriction of T' terms (cF:1396-1565, 2733-2810, 2857-2914, 3057-3132,
// 3169-3221).  A group of GL lanes owns one covered coarse element c and sums,
// in FP64, the contributions of the fine elements whose Gauss point 0 lies in
// c to its 8 nodes:
//   -w [ grad Nc . k_e grad T'(q) + Nc rho*cp_e dT'(q) / dt ]
// with Nc and grad Nc evaluated (per axis) in the Gauss point's own coarse
// element, as the reference's Shapes tensors are.
// ---------------------------------------------------------------------------

template <int GL>
__global__ void __launch_bounds__(256) k_restrict(GmPair p, const GmStepDesc* __restrict__ d,
                                                  DevMat m, RArgs r, double* __restrict__ part) {
    const int lane = threadIdx.x & (GL - 1);
    const int gid = (blockIdx.x * blockDim.x + threadIdx.x) / GL;
    const int cnx = p.ax[0].info[1], cny = p.ax[1].info[1], cnz = p.ax[2].info[1];
    const bool live = gid < cnx * cny * cnz;
    double acc[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) acc[a] = 0.0;
    if (live) {
        const int cx = gid % cnx, cy = (gid / cnx) % cny, cz = gid / (cnx * cny);
        const int ex0 = p.ax[0].info[2 + cx], ex1 = p.ax[0].info[2 + p.cmax[0] + cx];
        const int ey0 = p.ax[1].info[2 + cy], ey1 = p.ax[1].info[2 + p.cmax[1] + cy];
        const int ez0 = p.ax[2].info[2 + cz], ez1 = p.ax[2].info[2 + p.cmax[2] + cz];
        const int nfx = ex1 - ex0, nfy = ey1 - ey0, nfz = ez1 - ez0;
        const int cnt = nfx * nfy * nfz;
        const gm_grid& gf = d->g[p.fine];
        const gm_grid& gc = d->g[p.coarse];
        const int nx = gf.n[0], ny = gf.n[1];
        const int64_t nxy = (int64_t)nx * ny;
        const float ihf[3] = {1.0f / gf.h[0], 1.0f / gf.h[1], 1.0f / gf.h[2]};
        const float ihc[3] = {1.0f / gc.h[0], 1.0f / gc.h[1], 1.0f / gc.h[2]};
        float dt = r.dt;
        if (r.dt_mode == GM_DT_L1) dt = d->dt1;
        else if (r.dt_mode == GM_DT_L2) dt = d->dt2[r.dt_index];
        else if (r.dt_mode == GM_DT_ROW) dt = d->rows[r.dt_index][5];
        const float wflux = r.wf, wmass = (1.0f / dt) * r.wf;
        for (int k = lane; k < cnt; k += GL) {
            const int ex = ex0 + k % nfx, ey = ey0 + (k / nfx) % nfy, ez = ez0 + k / (nfx * nfy);
            const int64_t base = ex + (int64_t)nx * ey + nxy * ez;
            const int64_t off[8] = {0, 1, 1 + nx, nx, nxy, 1 + nxy, 1 + nx + nxy, nx + nxy};
            float kk = 0.0f, cc = 0.0f, tp[8], dtp[8];
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                const int64_t n = base + off[a];
                const float T = r.T[n];
                const float s1 = r.S1 ? (float)r.S1[n] : r.S1f[n];
                float kv, cv;
                bool b1, b2;
                props(m, T, s1, n < r.n_sub, kv, cv, b1, b2);
                kk += kv;
                cc += cv;
                tp[a] = r.Tp ? r.Tp[n] : 0.0f;
                dtp[a] = r.TpA ? (r.TpA[n] - r.TpB[n]) : 0.0f;
            }
            const float ke = kk * 0.125f, ce = cc * 0.125f;
            const float4 qx[2] = {p.ax[0].q[2 * ex], p.ax[0].q[2 * ex + 1]};
            const float4 qy[2] = {p.ax[1].q[2 * ey], p.ax[1].q[2 * ey + 1]};
            const float4 qz[2] = {p.ax[2].q[2 * ez], p.ax[2].q[2 * ez + 1]};
            float c8[8];
#pragma unroll
            for (int a = 0; a < 8; ++a) c8[a] = 0.0f;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int gx = ((q + 1) >> 1) & 1, gy = (q >> 1) & 1, gz = q >> 2;
                float gX = 0.0f, gY = 0.0f, gZ = 0.0f, ms = 0.0f;
#pragma unroll
                for (int b = 0; b < 8; ++b) {
                    const int bx = ((b + 1) >> 1) & 1, by = (b >> 1) & 1, bz = b >> 2;
                    const float nxf = (bx == gx) ? GA : GB, nyf = (by == gy) ? GA : GB,
                                nzf = (bz == gz) ? GA : GB;
                    const float sx = bx ? ihf[0] : -ihf[0], sy = by ? ihf[1] : -ihf[1],
                                sz = bz ? ihf[2] : -ihf[2];
                    gX += sx * nyf * nzf * tp[b];
                    gY += nxf * sy * nzf * tp[b];
                    gZ += nxf * nyf * sz * tp[b];
                    ms += nxf * nyf * nzf * dtp[b];
                }
                const float fX = ke * gX, fY = ke * gY, fZ = ke * gZ, mm = ce * ms;
                const float px[2] = {gx ? qx[1].y : qx[0].y, gx ? qx[1].z : qx[0].z};
                const float py[2] = {gy ? qy[1].y : qy[0].y, gy ? qy[1].z : qy[0].z};
                const float pz[2] = {gz ? qz[1].y : qz[0].y, gz ? qz[1].z : qz[0].z};
#pragma unroll
                for (int a = 0; a < 8; ++a) {
                    const int ax = ((a + 1) >> 1) & 1, ay = (a >> 1) & 1, az = a >> 2;
                    const float dpx = ax ? ihc[0] : -ihc[0], dpy = ay ? ihc[1] : -ihc[1],
                                dpz = az ? ihc[2] : -ihc[2];
                    const float Nc = px[ax] * py[ay] * pz[az];
                    const float fl = (dpx * py[ay] * pz[az]) * fX + (px[ax] * dpy * pz[az]) * fY +
                                     (px[ax] * py[ay] * dpz) * fZ;
                    c8[a] += -fl * wflux - Nc * mm * wmass;
                }
            }
#pragma unroll
            for (int a = 0; a < 8; ++a) acc[a] += (double)c8[a];
        }
    }
#pragma unroll
    for (int off = GL / 2; off > 0; off >>= 1)
#pragma unroll
        for (int a = 0; a < 8; ++a) acc[a] += __shfl_xor(acc[a], off, GL);
    if (live && lane == 0) {
#pragma unroll
        for (int a = 0; a < 8; ++a) part[(int64_t)gid * 8 + a] = acc[a];
    }
}

void launch_restrict(const GmPair& p, const GmStepDesc* d, const DevMat& m, const RArgs& r,
                     double* part, hipStream_t s) {
    // fine elements per coarse element decide the group width
    int per = 1;
    for (int dd = 0; dd < 3; ++dd) per *= (p.nef[dd] + p.cmax[dd] - 1) / p.cmax[dd];
    const int64_t groups = (int64_t)p.cmax[0] * p.cmax[1] * p.cmax[2];
    if (per >= 48) {
        const int64_t thr = groups * 64;
        hipLaunchKernelGGL(k_restrict<64>, dim3((thr + 255) / 256), dim3(256), 0, s, p, d, m, r,
                           part);
    } else if (per >= 12) {
        const int64_t thr = groups * 16;
        hipLaunchKernelGGL(k_restrict<16>, dim3((thr + 255) / 256), dim3(256), 0, s, p, d, m, r,
                           part);
    } else {
        const int64_t thr = groups * 4;
        hipLaunchKernelGGL(k_restrict<4>, dim3((thr + 255) / 256), dim3(256), 0, s, p, d, m, r,
                           part);
    }
}

// Sum the element partials of one (or two nested) pairs onto the coarse node
// box of the outer pair, plus (optionally) the coarse laser source
//   F_A = scale * sum_pos Gx Gy Gz
// of computeLevelSource on the node box of the source pair.  V has the
// layout [cmax+1]^3 of the outer pair.
__global__ void k_restrict_nodes(GmPair po, const double* __restrict__ parto, GmPair pi,
                                 const double* __restrict__ parti, int have_inner,
                                 const float* __restrict__ Vbase, float* __restrict__ V,
                                 GmPair pg, const float* __restrict__ G, int npos, int nmax,
                                 float scale) {
    const int mx = po.cmax[0] + 1, my = po.cmax[1] + 1, mz = po.cmax[2] + 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= mx * my * mz) return;
    const int bx = t % mx, by = (t / mx) % my, bz = t / (mx * my);
    const int A[3] = {po.ax[0].info[0] + bx, po.ax[1].info[0] + by, po.ax[2].info[0] + bz};
    double s = 0.0;
    const GmPair* pp[2] = {&po, &pi};
    const double* pa[2] = {parto, parti};
    for (int w = 0; w < 1 + have_inner; ++w) {
        const GmPair& P = *pp[w];
        const int lo[3] = {P.ax[0].info[0], P.ax[1].info[0], P.ax[2].info[0]};
        const int n[3] = {P.ax[0].info[1], P.ax[1].info[1], P.ax[2].info[1]};
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const int ax = ((a + 1) >> 1) & 1, ay = (a >> 1) & 1, az = a >> 2;
            const int cx = A[0] - ax - lo[0], cy = A[1] - ay - lo[1], cz = A[2] - az - lo[2];
            if (cx < 0 || cx >= n[0] || cy < 0 || cy >= n[1] || cz < 0 || cz >= n[2]) continue;
            s += pa[w][((int64_t)cx + (int64_t)n[0] * (cy + (int64_t)n[1] * cz)) * 8 + a];
        }
    }
    float v = (Vbase ? Vbase[t] : 0.0f) + (float)s;
    if (G) {
        const int gx = A[0] - pg.ax[0].info[0], gy = A[1] - pg.ax[1].info[0],
                  gz = A[2] - pg.ax[2].info[0];
        if (gx >= 0 && gx <= pg.ax[0].info[1] && gy >= 0 && gy <= pg.ax[1].info[1] && gz >= 0 &&
            gz <= pg.ax[2].info[1]) {
            float f = 0.0f;
            for (int q = 0; q < npos; ++q) {
                const float* Gq = G + q * 3 * nmax;
                f += Gq[gx] * Gq[nmax + gy] * Gq[2 * nmax + gz];
            }
            v += f * scale;
        }
    }
    V[t] = v;
}

void launch_restrict_nodes(const GmPair& outer, const double* part_outer, const GmPair* inner,
                           const double* part_inner, const float* Vbase, float* V,
                           const GmPair* gpair, const float* G, int npos, int nmax, float scale,
                           hipStream_t s) {
    const int n = (outer.cmax[0] + 1) * (outer.cmax[1] + 1) * (outer.cmax[2] + 1);
    GmPair pi = inner ? *inner : outer;
    GmPair pg = gpair ? *gpair : outer;
    hipLaunchKernelGGL(k_restrict_nodes, dim3((n + 255) / 256), dim3(256), 0, s, outer, part_outer,
                       pi, part_inner, inner ? 1 : 0, Vbase, V, pg, gpair ? G : nullptr, npos,
                       nmax, scale);
}

// ---------------------------------------------------------------------------
// Node-centric restriction, for pairs whose fine nodes lie on every coarse
// node of the overlap (integer spacing ratio, aligned windows).  There the
// coarse shape function N_A restricted to a fine element is the fine
// trilinear interpolant of its nodal values, and the 2x2x2 Gauss rule of the
// reference integrates its products exactly, so the restricted T' terms of
// cF:1396-1565 are
//   V_A = sum_i N_A(x_i) r_i,
//   r_i = -[ (K(k_e) T')_i + (M(rho*cp_e) dT')_i / dt ]
// with K and M the fine stiffness and consistent mass operators built from
// the element means (cF:625-626, 2773, 2895, 3106).  r is a 27-point stencil
// of the level step's form; N_A(x_i) factors per axis into (R - j)/R and j/R
// for a fine node j spacings past coarse node c (ratio R).  Sums over fine
// nodes run in FP64: per column over z in registers, then x and y in LDS,
// then one FP64 atomic add per coarse node and block.
// ---------------------------------------------------------------------------
#define RZ 4
#define NZA (RZ + 1)
#define NAXM (LX + 3)

__device__ __forceinline__ float stencil27(const float* kb, const float* ka, int q0,
                                           const float* Tm, const float* T0, const float* Tp,
                                           int hc, const float* co) {
    const float k000 = kb[q0], k100 = kb[q0 + 1], k010 = kb[q0 + QX], k110 = kb[q0 + QX + 1];
    const float k001 = ka[q0], k101 = ka[q0 + 1], k011 = ka[q0 + QX], k111 = ka[q0 + QX + 1];
    const float Pz00 = k000 + k001, Pz10 = k100 + k101, Pz01 = k010 + k011, Pz11 = k110 + k111;
    const float Py00 = k000 + k010, Py10 = k100 + k110, Py01 = k001 + k011, Py11 = k101 + k111;
    const float Px00 = k000 + k100, Px10 = k010 + k110, Px01 = k001 + k101, Px11 = k011 + k111;
    const float Fx0 = Pz00 + Pz01, Fx1 = Pz10 + Pz11;
    const float Fy0 = Pz00 + Pz10, Fy1 = Pz01 + Pz11;
    const float Fz0 = Py00 + Py10, Fz1 = Py01 + Py11;
    float s = co[0] * ((Fx0 + Fx1) * T0[hc]);
    s += co[1] * (Fx0 * T0[hc - 1] + Fx1 * T0[hc + 1]);
    s += co[2] * (Fy0 * T0[hc - PX] + Fy1 * T0[hc + PX]);
    s += co[4] * (Fz0 * Tm[hc] + Fz1 * Tp[hc]);
    s += co[3] * (Pz00 * T0[hc - 1 - PX] + Pz10 * T0[hc + 1 - PX] + Pz01 * T0[hc - 1 + PX] +
                  Pz11 * T0[hc + 1 + PX]);
    s += co[5] * (Py00 * Tm[hc - 1] + Py10 * Tm[hc + 1] + Py01 * Tp[hc - 1] + Py11 * Tp[hc + 1]);
    s += co[6] * (Px00 * Tm[hc - PX] + Px10 * Tm[hc + PX] + Px01 * Tp[hc - PX] +
                  Px11 * Tp[hc + PX]);
    s += co[7] * (k000 * Tm[hc - 1 - PX] + k100 * Tm[hc + 1 - PX] + k010 * Tm[hc - 1 + PX] +
                  k110 * Tm[hc + 1 + PX] + k001 * Tp[hc - 1 - PX] + k101 * Tp[hc + 1 - PX] +
                  k011 * Tp[hc - 1 + PX] + k111 * Tp[hc + 1 + PX]);
    return s;
}

template <bool KCM, bool FLUX, bool MASS>
__global__ void __launch_bounds__(256) k_restrict_kc(GmPair pp, const GmStepDesc* __restrict__ d,
                                                     const DevMat* __restrict__ m, RArgs r,
                                                     double* __restrict__ acc) {
    constexpr int NP = RZ + 2;
    __shared__ float sA[FLUX ? 3 : 1][PN];
    __shared__ float sB[MASS ? 3 : 1][PN];
    __shared__ float sKn[PN], sCn[PN];
    __shared__ float sSk[2][QN], sSc[2][QN];
    __shared__ float sKe[2][QN], sCe[2][QN];
    __shared__ float sTab[KCM ? 4 : PTW] __attribute__((aligned(16)));
    __shared__ double sZ[NZA][LY * LX];
    __shared__ double sX[NZA][LY][NAXM];

    const gm_grid gf = d->g[pp.fine];
    const gm_grid gc = d->g[pp.coarse];
    const int nx = gf.n[0], ny = gf.n[1], nz = gf.n[2];
    const int nxy = nx * ny;
    const int nn = nxy * nz;
    const int ntx = (nx + LX - 1) / LX, nty = (ny + LY - 1) / LY;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int ox = (t % ntx) * LX, oy = ((t / ntx) % nty) * LY, oz = (t / (ntx * nty)) * RZ;
    const int nex = nx - 1, ney = ny - 1, nez = nz - 1;
    const int tid = threadIdx.x;
    const int tx = tid % LX, ty = tid / LX;
    const int ix = ox + tx, iy = oy + ty;
    const bool col_ok = ix < nx && iy < ny;
    int R[3], off[3];
    double invR[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        // integer ratio and offset: one hardware reciprocal is exact enough to round
        const float ih = __builtin_amdgcn_rcpf(gf.h[a]);
        R[a] = (int)rintf(gc.h[a] * ih);
        off[a] = (int)rintf((gf.x0[a] - gc.x0[a]) * ih);
        invR[a] = 1.0 / (double)R[a];
    }

    int eoff[2];
    bool eok[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const int l = tid + 256 * e;
        const int jx = ox + l % PX - 1, jy = oy + l / PX - 1;
        eok[e] = l < PN && jx >= 0 && jx < nx && jy >= 0 && jy < ny;
        eoff[e] = jx + nx * jy;
    }
    // all global loads of the block first
    float rA[NP][2], rB[NP][2], rK[NP][2], rC[NP][2];
#pragma unroll
    for (int pl = 0; pl < NP; ++pl) {
        const int p = oz - 1 + pl;
        const bool pin = p >= 0 && p < nz;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            rA[pl][e] = rB[pl][e] = rK[pl][e] = rC[pl][e] = 0.0f;
            if (pin && eok[e]) {
                const int idx = eoff[e] + nxy * p;
                if (FLUX) rA[pl][e] = r.Tp[idx];
                if (MASS) rB[pl][e] = r.TpA[idx] - r.TpB[idx];
                if (KCM) {
                    rK[pl][e] = r.KC[idx];
                    rC[pl][e] = r.KC[nn + idx];
                } else {
                    rK[pl][e] = r.T[idx];
                    rC[pl][e] = (float)r.S1[idx];
                }
            }
        }
    }
    float dt = r.dt;
    if (r.dt_mode == GM_DT_L1) dt = d->dt1;
    else if (r.dt_mode == GM_DT_L2) dt = d->dt2[r.dt_index];
    else if (r.dt_mode == GM_DT_ROW) dt = d->rows[r.dt_index][5];
    const float idt = 1.0f / dt;
    PropC pc;
    if (!KCM) {
        pc = load_pc(m);
        stage_tables(m, sTab, tid, 256);
    }
    int qh[2];
    bool qok[2], qxy[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const int l = tid + 256 * e;
        const int qx = l % QX, qy = l / QX;
        qh[e] = qx + PX * qy;
        const int ex = ox + qx - 1, ey = oy + qy - 1;
        qok[e] = l < QN;
        qxy[e] = l < QN && ex >= 0 && ex < nex && ey >= 0 && ey < ney;
    }
    const int hc = (tx + 1) + PX * (ty + 1);
    const int q0 = tx + QX * ty;
    const int AzB = (oz + off[2]) / R[2];
    double zacc[NZA];
#pragma unroll
    for (int a = 0; a < NZA; ++a) zacc[a] = 0.0;

#pragma unroll
    for (int pl = 0; pl < NP; ++pl) {
        const int p = oz - 1 + pl;
        lds_barrier();
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int l = tid + 256 * e;
            if (l < PN) {
                if (FLUX) sA[pl % 3][l] = rA[pl][e];
                if (MASS) sB[pl % 3][l] = rB[pl][e];
                float kv = rK[pl][e], cv = rC[pl][e];
                if (!KCM && eok[e] && p >= 0 && p < nz) {
                    bool s1;
                    props_c(pc, sTab, rK[pl][e], rC[pl][e] > 0.5f, eoff[e] + nxy * p < r.n_sub,
                            kv, cv, s1);
                }
                sKn[l] = kv;
                sCn[l] = cv;
            }
        }
        lds_barrier();
        const bool ezok = p >= 1 && p - 1 < nez;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            if (qok[e]) {
                const int h0 = qh[e], l = tid + 256 * e;
                const float Sk = (sKn[h0] + sKn[h0 + 1]) + (sKn[h0 + PX] + sKn[h0 + PX + 1]);
                const float Sc = (sCn[h0] + sCn[h0 + 1]) + (sCn[h0 + PX] + sCn[h0 + PX + 1]);
                const bool ok = qxy[e] && ezok;
                sKe[pl & 1][l] = ok ? (sSk[(pl + 1) & 1][l] + Sk) * 0.125f : 0.0f;
                sCe[pl & 1][l] = ok ? (sSc[(pl + 1) & 1][l] + Sc) * 0.125f : 0.0f;
                sSk[pl & 1][l] = Sk;
                sSc[pl & 1][l] = Sc;
            }
        }
        lds_barrier();
        if (pl < 2) continue;
        const int zo = oz + pl - 2;
        if (zo >= nz || !col_ok) continue;
        float rr = 0.0f;
        if (FLUX)
            rr -= stencil27(sKe[(pl + 1) & 1], sKe[pl & 1], q0, sA[(pl + 1) % 3], sA[(pl + 2) % 3],
                            sA[pl % 3], hc, r.kcs);
        if (MASS)
            rr -= idt * stencil27(sCe[(pl + 1) & 1], sCe[pl & 1], q0, sB[(pl + 1) % 3],
                                  sB[(pl + 2) % 3], sB[pl % 3], hc, r.mcs);
        const int u = zo + off[2];
        const int c = u / R[2];
        const int j = u - c * R[2];
        const int lz = c - AzB;
        const double v = (double)rr;
        const double v1 = (double)j * invR[2] * v, v0 = v - v1;
#pragma unroll
        for (int a = 0; a < NZA; ++a) zacc[a] += (a == lz ? v0 : (a == lz + 1 ? v1 : 0.0));
    }
#pragma unroll
    for (int a = 0; a < NZA; ++a) sZ[a][tid] = zacc[a];
    lds_barrier();
    // x: coarse node Ax gets sum over the tile's fine columns with weight 1 - |u - Ax R| / R
    const int AxB = (ox + off[0]) / R[0], AyB = (oy + off[1]) / R[1];
    const int xl = min(ox + LX, nx) - 1, yl = min(oy + LY, ny) - 1, zl = min(oz + RZ, nz) - 1;
    const int NAX = (xl + off[0]) / R[0] - AxB + 2;
    const int NAY = (yl + off[1]) / R[1] - AyB + 2;
    const int NAZ = (zl + off[2]) / R[2] - AzB + 2;
    for (int cI = tid; cI < NAZ * LY * NAX; cI += 256) {
        const int az = cI / (LY * NAX), rem = cI % (LY * NAX);
        const int yy = rem / NAX, ax = rem % NAX;
        const int Au = (AxB + ax) * R[0] - off[0];  // fine index of coarse node Ax
        double sum = 0.0;
        const int lo = max(Au - R[0] + 1, ox), hi = min(Au + R[0] - 1, xl);
        for (int i = lo; i <= hi; ++i) {
            const int dd = i - Au;
            sum += (double)(R[0] - (dd < 0 ? -dd : dd)) * invR[0] * sZ[az][yy * LX + i - ox];
        }
        sX[az][yy][ax] = sum;
    }
    lds_barrier();
    const int lo0 = __builtin_amdgcn_readfirstlane(pp.ax[0].info[0]);
    const int lo1 = __builtin_amdgcn_readfirstlane(pp.ax[1].info[0]);
    const int lo2 = __builtin_amdgcn_readfirstlane(pp.ax[2].info[0]);
    const int nb0 = __builtin_amdgcn_readfirstlane(pp.ax[0].info[1]) + 1;
    const int nb1 = __builtin_amdgcn_readfirstlane(pp.ax[1].info[1]) + 1;
    const int nb2 = __builtin_amdgcn_readfirstlane(pp.ax[2].info[1]) + 1;
    const int st0 = pp.cmax[0] + 1, st1 = pp.cmax[1] + 1;
    for (int cI = tid; cI < NAZ * NAY * NAX; cI += 256) {
        const int az = cI / (NAY * NAX), rem = cI % (NAY * NAX);
        const int ay = rem / NAX, ax = rem % NAX;
        const int Au = (AyB + ay) * R[1] - off[1];
        double sum = 0.0;
        const int lo = max(Au - R[1] + 1, oy), hi = min(Au + R[1] - 1, yl);
        for (int i = lo; i <= hi; ++i) {
            const int dd = i - Au;
            sum += (double)(R[1] - (dd < 0 ? -dd : dd)) * invR[1] * sX[az][i - oy][ax];
        }
        const int bx = AxB + ax - lo0, by = AyB + ay - lo1, bz = AzB + az - lo2;
        if (sum != 0.0 && bx >= 0 && bx < nb0 && by >= 0 && by < nb1 && bz >= 0 && bz < nb2)
            unsafeAtomicAdd(acc + (bx + (int64_t)st0 * (by + (int64_t)st1 * bz)), sum);
    }
}

// Node-centric restriction with the carried k, cp of the fine level, one
// wave per brick like the level step.  A wave owns an x-strip of W fine
// columns (W a multiple of the x ratio R, starting on a coarse node), one
// fine row and ZC fine planes; lane l holds x = X - 1 + l.  Per fine node
// the residual r of the flux and mass terms is the 27-point stencil of the
// level step; the z weights (uniform in the wave) fold it into per-lane FP64
// sums per coarse z node, then the lane on each coarse node gathers the
// x weights (R - |d|)/R of its 2R-1 neighbours through DPP shifts, applies
// the y weight and adds into the pair's FP64 node box.
__device__ __forceinline__ double dprev(double v) {  // lane l-1 (0 in lane 0)
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)u, 0x138, 0xf, 0xf, true);
    const uint32_t hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), 0x138, 0xf, 0xf, true);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ double dnext(double v) {  // lane l+1 (0 in lane 63)
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)u, 0x130, 0xf, 0xf, true);
    const uint32_t hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), 0x130, 0xf, 0xf, true);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// 27-point stencil of the level-step form at the lane's node of window
// plane i+1, row 1: element values E (planes i, i+1; rows 0, 1 = y-1, y
// sides), field U over planes i..i+2, rows 0..2, pattern coefficients co.
template <int NR>
__device__ __forceinline__ float stencil_dpp(const float (*E)[2], const float (*U)[NR], int i,
                                             const float* co) {
    const float k100 = E[i][0], k110 = E[i][1], k101 = E[i + 1][0], k111 = E[i + 1][1];
    const float k000 = lane_prev(k100), k010 = lane_prev(k110);
    const float k001 = lane_prev(k101), k011 = lane_prev(k111);
    const float Ti = U[i + 1][1];
    const float Pz00 = k000 + k001, Pz10 = k100 + k101, Pz01 = k010 + k011, Pz11 = k110 + k111;
    const float Py00 = k000 + k010, Py10 = k100 + k110, Py01 = k001 + k011, Py11 = k101 + k111;
    const float Px00 = k000 + k100, Px10 = k010 + k110, Px01 = k001 + k101, Px11 = k011 + k111;
    const float Fx0 = Pz00 + Pz01, Fx1 = Pz10 + Pz11;
    const float Fy0 = Pz00 + Pz10, Fy1 = Pz01 + Pz11;
    const float Fz0 = Py00 + Py10, Fz1 = Py01 + Py11;
    float s = co[0] * ((Fx0 + Fx1) * Ti);
    s += co[1] * (Fx0 * lane_prev(Ti) + Fx1 * lane_next(Ti));
    s += co[2] * (Fy0 * U[i + 1][0] + Fy1 * U[i + 1][2]);
    s += co[4] * (Fz0 * U[i][1] + Fz1 * U[i + 2][1]);
    s += co[3] * (Pz00 * lane_prev(U[i + 1][0]) + Pz10 * lane_next(U[i + 1][0]) +
                  Pz01 * lane_prev(U[i + 1][2]) + Pz11 * lane_next(U[i + 1][2]));
    s += co[5] * (Py00 * lane_prev(U[i][1]) + Py10 * lane_next(U[i][1]) +
                  Py01 * lane_prev(U[i + 2][1]) + Py11 * lane_next(U[i + 2][1]));
    s += co[6] * (Px00 * U[i][0] + Px10 * U[i][2] + Px01 * U[i + 2][0] + Px11 * U[i + 2][2]);
    s += co[7] * (k000 * lane_prev(U[i][0]) + k100 * lane_next(U[i][0]) +
                  k010 * lane_prev(U[i][2]) + k110 * lane_next(U[i][2]) +
                  k001 * lane_prev(U[i + 2][0]) + k101 * lane_next(U[i + 2][0]) +
                  k011 * lane_prev(U[i + 2][2]) + k111 * lane_next(U[i + 2][2]));
    return s;
}

// One wave's brick t of the node-centric restriction of pair pp (the body of
// k_restrict_dpp and of its batched form).  Values read through d, pp and r
// are the same in every lane; they are moved to scalar registers where they
// size buffers or offsets, since in the batched launch the compiler cannot see
// that.
template <bool FLUX, bool MASS, int ZC>
__device__ __forceinline__ void restrict_dpp_brick(const GmPair& pp,
                                                   const GmStepDesc* __restrict__ d,
                                                   const RArgs& r, double* __restrict__ acc,
                                                   int t) {
    constexpr int NZC = ZC + 1;  // coarse z nodes a brick can touch (ratio 1)
    const gm_grid& gf = d->g[pp.fine];
    const gm_grid& gc = d->g[pp.coarse];
    const int nx = __builtin_amdgcn_readfirstlane(gf.n[0]);
    const int ny = __builtin_amdgcn_readfirstlane(gf.n[1]);
    const int nz = __builtin_amdgcn_readfirstlane(gf.n[2]);
    const int nxy = nx * ny;
    const int nn = nxy * nz;
    const int nex = nx - 1, ney = ny - 1, nez = nz - 1;
    // ratios, strip width and phase from the host (the launch grid follows
    // them); fine-origin offsets in fine spacings from the descriptor
    const int R[3] = {r.R0, r.R1, r.R2};
    const float iRf[3] = {__builtin_amdgcn_rcpf((float)r.R0), __builtin_amdgcn_rcpf((float)r.R1),
                          __builtin_amdgcn_rcpf((float)r.R2)};
    int off[3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
        off[a] = __builtin_amdgcn_readfirstlane(
            (int)rintf((gf.x0[a] - gc.x0[a]) * __builtin_amdgcn_rcpf(gf.h[a])));
    const int W = r.W, a0 = r.a0;  // a0: first fine x on a coarse node
    const int s0 = a0 > 0 ? 0 : 1;
    const int nzc = (nz + ZC - 1) / ZC;
    const int rest = fdiv_small(t, __builtin_amdgcn_rcpf((float)nzc));
    const int zc = t - rest * nzc;
    const int st = fdiv_small(rest, __builtin_amdgcn_rcpf((float)ny));
    const int y = rest - st * ny;
    const int z0 = zc * ZC;
    const int X = a0 + (st + s0 - 1) * W;  // strip start (a coarse node)
    const int lane = threadIdx.x;
    const int x = X - 1 + lane;
    const bool xin = x >= 0 && x < nx;
    const bool ein = x >= 0 && x < nex;
    const bool own = lane >= 1 && lane <= W && xin;
    const uint32_t nb4 = (uint32_t)nn * 4u;
    const int base = x + nx * y + nxy * z0;

    // ---- loads: planes z0-1 .. z0+ZC, rows y-1 .. y+1 ------------------------
    const auto rKC = buf_rsrc(r.KC, 2u * nb4);
    const auto rP = buf_rsrc(FLUX ? r.Tp : r.KC, nb4);
    const auto rA = buf_rsrc(MASS ? r.TpA : r.KC, nb4), rB = buf_rsrc(MASS ? r.TpB : r.KC, nb4);
    float Pw[ZC + 2][3], Dw[ZC + 2][3], Kw[ZC + 2][3], Cw[ZC + 2][3];
#pragma unroll
    for (int p = 0; p < ZC + 2; ++p) {
        const int zz = z0 - 1 + p;
#pragma unroll
        for (int rr = 0; rr < 3; ++rr) {
            const int yy = y - 1 + rr;
            const bool in = xin && yy >= 0 && yy < ny && zz >= 0 && zz < nz;
            const uint32_t vo = in ? 4u * (uint32_t)(base + nx * (rr - 1) + nxy * (p - 1)) : GM_OOB;
            Kw[p][rr] = buf_ld(rKC, vo);
            Cw[p][rr] = buf_ld(rKC, vo, nb4);
            Pw[p][rr] = FLUX ? buf_ld(rP, vo) : 0.0f;
            Dw[p][rr] = MASS ? buf_ld(rA, vo) - buf_ld(rB, vo) : 0.0f;
        }
    }
    float dt = r.dt;
    if (r.dt_mode == GM_DT_L1) dt = d->dt1;
    else if (r.dt_mode == GM_DT_L2) dt = d->dt2[r.dt_index];
    else if (r.dt_mode == GM_DT_ROW) dt = d->rows[r.dt_index][5];
    const float idt = 1.0f / dt;

    // ---- element means of k and cp (as the level step) ---------------------
    float Qk[ZC + 2][2], Qc[ZC + 2][2];
#pragma unroll
    for (int p = 0; p < ZC + 2; ++p) {
#pragma unroll
        for (int rr = 0; rr < 2; ++rr) {
            const int yy = y - 1 + rr;
            const bool ok = ein && yy >= 0 && yy < ney;
            const float ak = Kw[p][rr] + Kw[p][rr + 1];
            const float ac = Cw[p][rr] + Cw[p][rr + 1];
            const float qk = ak + lane_next(ak);
            const float qc = ac + lane_next(ac);
            Qk[p][rr] = ok ? qk : 0.0f;
            Qc[p][rr] = ok ? qc : 0.0f;
        }
    }
    float Ek[ZC + 1][2], Ec[ZC + 1][2];
#pragma unroll
    for (int p = 0; p < ZC + 1; ++p) {
        const int ez = z0 - 1 + p;
        const float f = (ez >= 0 && ez < nez) ? 0.125f : 0.0f;
#pragma unroll
        for (int rr = 0; rr < 2; ++rr) {
            Ek[p][rr] = (Qk[p][rr] + Qk[p + 1][rr]) * f;
            Ec[p][rr] = (Qc[p][rr] + Qc[p + 1][rr]) * f;
        }
    }

    // ---- residuals, folded over z into coarse z nodes -------------------------
    const int czb = fdiv_small(z0 + off[2], iRf[2]);
    const double invR2 = r.invR[2];
    double zacc[NZC];
#pragma unroll
    for (int b = 0; b < NZC; ++b) zacc[b] = 0.0;
#pragma unroll
    for (int i = 0; i < ZC; ++i) {
        const int zz = z0 + i;
        float rv = 0.0f;
        if (FLUX) rv -= stencil_dpp<3>(Ek, Pw, i, r.kcs);
        if (MASS) rv -= idt * stencil_dpp<3>(Ec, Dw, i, r.mcs);
        const double v = (own && zz < nz) ? (double)rv : 0.0;
        const int uz = zz + off[2];
        const int cz = fdiv_small(uz, iRf[2]);
        const int jz = uz - cz * R[2];
        const int lz = cz - czb;
        const double w1 = (double)jz * invR2, w0 = 1.0 - w1;
#pragma unroll
        for (int b = 0; b < NZC; ++b) {
            const double w = b == lz ? w0 : (b == lz + 1 ? w1 : 0.0);
            zacc[b] += w * v;
        }
    }

    // ---- x gather on the coarse-node lanes, y weight, box add -----------------
    const int uy = y + off[1];
    const int cy = fdiv_small(uy, iRf[1]);
    const int jy = uy - cy * R[1];
    const double wy1 = (double)jy * r.invR[1], wy0 = 1.0 - wy1;
    const double invR0 = r.invR[0];
    // coarse-node lanes, also those just outside the fine mesh whose support
    // reaches into it (x + off is a multiple of R there)
    const int lq = fdiv_small(lane + R[0] - 1, iRf[0]) - 1;  // floor((lane - 1) / R)
    const bool owner = lane >= 1 && lane <= W + 1 && lane - 1 == lq * R[0];
    const int cx = fdiv_small(x + off[0] + 128 * R[0], iRf[0]) - 128;
    const int lo0 = __builtin_amdgcn_readfirstlane(pp.ax[0].info[0]);
    const int lo1 = __builtin_amdgcn_readfirstlane(pp.ax[1].info[0]);
    const int lo2 = __builtin_amdgcn_readfirstlane(pp.ax[2].info[0]);
    const int nb0 = __builtin_amdgcn_readfirstlane(pp.ax[0].info[1]) + 1;
    const int nb1 = __builtin_amdgcn_readfirstlane(pp.ax[1].info[1]) + 1;
    const int nb2 = __builtin_amdgcn_readfirstlane(pp.ax[2].info[1]) + 1;
    const int st0 = pp.cmax[0] + 1, st1 = pp.cmax[1] + 1;
#pragma unroll
    for (int b = 0; b < NZC; ++b) {
        const int bz = czb + b - lo2;
        if (b > fdiv_small(ZC + R[2] - 2, iRf[2]) + 1 || bz < 0 || bz >= nb2) continue;  // uniform
        double V = zacc[b], vp = zacc[b], vn = zacc[b];
        for (int dd = 1; dd < R[0]; ++dd) {
            vp = dprev(vp);
            vn = dnext(vn);
            V += (double)(R[0] - dd) * invR0 * (vp + vn);
        }
        const int bx = cx - lo0;
        if (owner && V != 0.0 && bx >= 0 && bx < nb0) {
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const int by = cy + a - lo1;
                const double w = a ? wy1 : wy0;
                if (w != 0.0 && by >= 0 && by < nb1)
                    unsafeAtomicAdd(acc + (bx + (int64_t)st0 * (by + (int64_t)st1 * bz)), w * V);
            }
        }
    }
}

template <bool FLUX, bool MASS, int ZC>
__global__ void __launch_bounds__(64) k_restrict_dpp(GmPair pp, const GmStepDesc* __restrict__ d,
                                                     RArgs r, double* __restrict__ acc) {
    restrict_dpp_brick<FLUX, MASS, ZC>(pp, d, r, acc, xcd_tile(blockIdx.x, gridDim.x));
}

// The restriction of many engines' pairs in one launch (items as the single
// launch's arguments, node-aligned pairs with carried k, cp): bstart[i] is
// the first block of item i.
template <bool FLUX, bool MASS, int ZC>
__global__ void __launch_bounds__(64) k_restrict_dpp_batch(const RItem* __restrict__ it,
                                                           const int* __restrict__ bstart,
                                                           int n) {
    const int b = xcd_tile(blockIdx.x, gridDim.x);
    int b0;
    const int lo = batch_item_of(bstart, n, b, &b0);
    const RItem& x = it[lo];
    restrict_dpp_brick<FLUX, MASS, ZC>(x.p, x.d, x.r, x.acc, b - b0);
}

// bricks of the node-aligned restriction of pair p with arguments r
static int restrict_dpp_nbd(const GmPair& p, const RArgs& r, int ZC) {
    const int nx = p.nef[0] + 1, ny = p.nef[1] + 1, nz = p.nef[2] + 1;
    const int nst = (nx - r.a0 + r.W - 1) / r.W + (r.a0 > 0 ? 1 : 0);
    return nst * ny * ((nz + ZC - 1) / ZC);
}

bool restrict_batchable(const RArgs& r) { return r.KC && r.dpp_ok && (r.Tp || r.TpA); }

int restrict_batch_blocks(const GmPair& p, const RArgs& r) { return restrict_dpp_nbd(p, r, 4); }

void launch_restrict_batch(const RItem* it, const int* bstart, int n, int nblocks, bool flux,
                           bool mass, hipStream_t s) {
    if (n <= 0 || nblocks <= 0) return;
    constexpr int ZC = 4;
    if (flux && mass)
        hipLaunchKernelGGL((k_restrict_dpp_batch<true, true, ZC>), dim3(nblocks), dim3(64), 0, s,
                           it, bstart, n);
    else if (flux)
        hipLaunchKernelGGL((k_restrict_dpp_batch<true, false, ZC>), dim3(nblocks), dim3(64), 0, s,
                           it, bstart, n);
    else
        hipLaunchKernelGGL((k_restrict_dpp_batch<false, true, ZC>), dim3(nblocks), dim3(64), 0, s,
                           it, bstart, n);
}

void launch_restrict_fast(const GmPair& p, const GmStepDesc* d, const DevMat* m, const RArgs& r,
                          double* acc, hipStream_t s) {
    // grid of the fine level: the pair's element counts + 1
    const int nx = p.nef[0] + 1, ny = p.nef[1] + 1, nz = p.nef[2] + 1;
    const int nb = ((nx + LX - 1) / LX) * ((ny + LY - 1) / LY) * ((nz + RZ - 1) / RZ);
    const bool fl = r.Tp != nullptr, ms = r.TpA != nullptr;
    if (r.KC && r.dpp_ok) {
        // x strips of W = R * floor(61 / R) columns starting on coarse nodes,
        // one per fine row and ZC planes
        constexpr int ZC = 4;
        const int nbd = restrict_dpp_nbd(p, r, ZC);
        if (fl && ms)
            hipLaunchKernelGGL((k_restrict_dpp<true, true, ZC>), dim3(nbd), dim3(64), 0, s, p, d, r,
                               acc);
        else if (fl)
            hipLaunchKernelGGL((k_restrict_dpp<true, false, ZC>), dim3(nbd), dim3(64), 0, s, p, d,
                               r, acc);
        else
            hipLaunchKernelGGL((k_restrict_dpp<false, true, ZC>), dim3(nbd), dim3(64), 0, s, p, d,
                               r, acc);
    } else if (r.KC) {
        if (fl && ms)
            hipLaunchKernelGGL((k_restrict_kc<true, true, true>), dim3(nb), dim3(256), 0, s, p, d,
                               m, r, acc);
        else if (fl)
            hipLaunchKernelGGL((k_restrict_kc<true, true, false>), dim3(nb), dim3(256), 0, s, p,
                               d, m, r, acc);
        else
            hipLaunchKernelGGL((k_restrict_kc<true, false, true>), dim3(nb), dim3(256), 0, s, p,
                               d, m, r, acc);
    } else {
        if (fl && ms)
            hipLaunchKernelGGL((k_restrict_kc<false, true, true>), dim3(nb), dim3(256), 0, s, p,
                               d, m, r, acc);
        else if (fl)
            hipLaunchKernelGGL((k_restrict_kc<false, true, false>), dim3(nb), dim3(256), 0, s, p,
                               d, m, r, acc);
        else
            hipLaunchKernelGGL((k_restrict_kc<false, false, true>), dim3(nb), dim3(256), 0, s, p,
                               d, m, r, acc);
    }
}

// V = Vbase + outer node box (+ inner node box) (+ coarse source), then the
// node boxes are cleared for the next accumulation.
__global__ void k_restrict_nodes_acc(GmPair po, double* __restrict__ acco, GmPair pi,
                                     double* __restrict__ acci, int have_inner,
                                     const float* __restrict__ Vbase, float* __restrict__ V,
                                     GmPair pg, const float* __restrict__ F, float scale) {
    const int mx = po.cmax[0] + 1, my = po.cmax[1] + 1, mz = po.cmax[2] + 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= mx * my * mz) return;
    const int bx = t % mx, by = (t / mx) % my, bz = t / (mx * my);
    const int A[3] = {po.ax[0].info[0] + bx, po.ax[1].info[0] + by, po.ax[2].info[0] + bz};
    double s = acco[t];
    acco[t] = 0.0;
    if (have_inner) {
        const int ix = A[0] - pi.ax[0].info[0], iy = A[1] - pi.ax[1].info[0],
                  iz = A[2] - pi.ax[2].info[0];
        if (ix >= 0 && ix <= pi.ax[0].info[1] && iy >= 0 && iy <= pi.ax[1].info[1] && iz >= 0 &&
            iz <= pi.ax[2].info[1]) {
            const int64_t k = ix + (int64_t)(pi.cmax[0] + 1) * (iy + (int64_t)(pi.cmax[1] + 1) * iz);
            s += acci[k];
        }
    }
    float v = (Vbase ? Vbase[t] : 0.0f) + (float)s;
    if (F) {
        // coarse source summed over the positions on the node box of pg
        const int gx = A[0] - pg.ax[0].info[0], gy = A[1] - pg.ax[1].info[0],
                  gz = A[2] - pg.ax[2].info[0];
        if (gx >= 0 && gx <= pg.ax[0].info[1] && gy >= 0 && gy <= pg.ax[1].info[1] && gz >= 0 &&
            gz <= pg.ax[2].info[1])
            v += F[gx + (pg.cmax[0] + 1) * (gy + (pg.cmax[1] + 1) * gz)] * scale;
    }
    V[t] = v;
}

// Coarse laser source on the node box of pair p (cF:2667-2730 summed over
// positions): F[k][t] = sum over the npos positions of group k of the
// per-axis table products, for every node t of the box (zero outside the
// covered range).  WAVE: one wave per node, lanes over the positions.
template <bool WAVE>
__global__ void k_src_box(GmPair p, const float* __restrict__ G, int nk, int npos, int nmax,
                          float* __restrict__ F) {
    const int mx = p.cmax[0] + 1, my = p.cmax[1] + 1, mz = p.cmax[2] + 1;
    const int nbox = mx * my * mz;
    const int o = WAVE ? blockIdx.x : blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= nk * nbox) return;
    const int k = o / nbox, t = o - k * nbox;
    const int gx = t % mx, gy = (t / mx) % my, gz = t / (mx * my);
    const bool in = gx <= p.ax[0].info[1] && gy <= p.ax[1].info[1] && gz <= p.ax[2].info[1];
    float f = 0.0f;
    if (in) {
        const float* Gk = G + (size_t)k * npos * 3 * nmax;
        if (WAVE) {
            for (int q = threadIdx.x; q < npos; q += 64) {
                const float* Gq = Gk + q * 3 * nmax;
                f += Gq[gx] * Gq[nmax + gy] * Gq[2 * nmax + gz];
            }
#pragma unroll
            for (int sh = 32; sh >= 1; sh >>= 1) f += __shfl_xor(f, sh, 64);
        } else {
            for (int q = 0; q < npos; ++q) {
                const float* Gq = Gk + q * 3 * nmax;
                f += Gq[gx] * Gq[nmax + gy] * Gq[2 * nmax + gz];
            }
        }
    }
    if (!WAVE || threadIdx.x == 0) F[o] = f;
}

void launch_src_box(const GmPair& p, const float* G, int nk, int npos, int nmax, float* F,
                    hipStream_t s) {
    const int n = nk * (p.cmax[0] + 1) * (p.cmax[1] + 1) * (p.cmax[2] + 1);
    if (npos > 16)
        hipLaunchKernelGGL(k_src_box<true>, dim3(n), dim3(64), 0, s, p, G, nk, npos, nmax, F);
    else
        hipLaunchKernelGGL(k_src_box<false>, dim3((n + 255) / 256), dim3(256), 0, s, p, G, nk,
                           npos, nmax, F);
}

// the inner node box is cleared by a second pass (its layout differs)
__global__ void k_clear_f64(double* __restrict__ a, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

void launch_restrict_nodes_acc(const GmPair& outer, double* acc_outer, const GmPair* inner,
                               double* acc_inner, const float* Vbase, float* V,
                               const GmPair* gpair, const float* F, float scale, hipStream_t s) {
    const int n = (outer.cmax[0] + 1) * (outer.cmax[1] + 1) * (outer.cmax[2] + 1);
    GmPair pi = inner ? *inner : outer;
    GmPair pg = gpair ? *gpair : outer;
    hipLaunchKernelGGL(k_restrict_nodes_acc, dim3((n + 255) / 256), dim3(256), 0, s, outer,
                       acc_outer, pi, acc_inner, inner ? 1 : 0, Vbase, V, pg, gpair ? F : nullptr,
                       scale);
    if (inner) {
        const int64_t ni = (int64_t)(inner->cmax[0] + 1) * (inner->cmax[1] + 1) *
                           (inner->cmax[2] + 1);
        hipLaunchKernelGGL(k_clear_f64, dim3((unsigned)((ni + 255) / 256)), dim3(256), 0, s,
                           acc_inner, ni);
    }
}

// ---------------------------------------------------------------------------
// Gathers: interpolation onto grids, injection, T', window move
// ---------------------------------------------------------------------------

__device__ __forceinline__ const gm_grid& pick_grid(const GmStepDesc* d, int mode,
                                                    const gm_grid& host) {
    if (mode == 10) return d->ov3_xyz;
    if (mode == 11) return d->ov2_xyz;
    if (mode >= 1 && mode <= 3) return d->g[mode];
    return host;
}

__global__ void k_interp_grid(const GmStepDesc* __restrict__ d, int src_mode, gm_grid src_h,
                              const float* __restrict__ u, const uint8_t* __restrict__ u8,
                              int tgt_mode, gm_grid tgt_h, float* __restrict__ out) {
    const gm_grid& tg = pick_grid(d, tgt_mode, tgt_h);
    const gm_grid& sg = pick_grid(d, src_mode, src_h);
    const int64_t n = (int64_t)tg.n[0] * tg.n[1] * tg.n[2];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int ix, iy, iz;
    dec3(t, tg.n[0], tg.n[1], ix, iy, iz);
    float N[8];
    int64_t nd[8];
    interp8(sg, coordf(tg, 0, ix), coordf(tg, 1, iy), coordf(tg, 2, iz), N, nd);
    float s = 0.0f;
#pragma unroll
    for (int a = 0; a < 8; ++a) s += N[a] * (u ? u[nd[a]] : (float)u8[nd[a]]);
    out[t] = s;
}

void launch_interp_grid(const GmStepDesc* d, int src_level, const gm_grid* src_host,
                        const float* u, const uint8_t* u8, int tgt_mode, const gm_grid* tgt_host,
                        int64_t n_target, float* out, hipStream_t s) {
    if (n_target <= 0) return;
    gm_grid sh = src_host ? *src_host : gm_grid{};
    gm_grid th = tgt_host ? *tgt_host : gm_grid{};
    hipLaunchKernelGGL(k_interp_grid, dim3((n_target + 255) / 256), dim3(256), 0, s, d,
                       src_host ? -1 : src_level, sh, u, u8, tgt_host ? -1 : tgt_mode, th, out);
}

// CoarseT[overlap box] <- fine field at the overlap coordinates (cF:2086-2094)
__global__ void k_inject(const GmStepDesc* __restrict__ d, int fine_level, int coarse_level,
                         const float* __restrict__ fT, float* __restrict__ cT) {
    const GmBox& b = fine_level == 3 ? d->ov3 : d->ov2;
    const gm_grid& xyz = fine_level == 3 ? d->ov3_xyz : d->ov2_xyz;
    const int n = b.n[0] * b.n[1] * b.n[2];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int ix, iy, iz;
    dec3(t, b.n[0], b.n[1], ix, iy, iz);
    const float v = interp_f(d->g[fine_level], fT, coordf(xyz, 0, ix), coordf(xyz, 1, iy),
                             coordf(xyz, 2, iz));
    const gm_grid& cg = d->g[coarse_level];
    const int cx = b.lo[0] + ix, cy = b.lo[1] + iy, cz = b.lo[2] + iz;
    if (cx < 0 || cx >= cg.n[0] || cy < 0 || cy >= cg.n[1] || cz < 0 || cz >= cg.n[2]) return;
    cT[nid(cx, cy, cz, cg.n[0], cg.n[1])] = v;
}

void launch_inject(const GmStepDesc* d, int fine_level, int coarse_level, const float* fineT,
                   float* coarseT, int64_t n_box, hipStream_t s) {
    if (n_box <= 0) return;
    hipLaunchKernelGGL(k_inject, dim3((n_box + 255) / 256), dim3(256), 0, s, d, fine_level,
                       coarse_level, fineT, coarseT);
}

// T' = fine - prolong(coarse) at every fine node (cF:2097)
__global__ void k_tprime(const GmStepDesc* __restrict__ d, int fine_level, int coarse_level,
                         const float* __restrict__ fT, const float* __restrict__ cT,
                         float* __restrict__ Tp) {
    const gm_grid& fg = d->g[fine_level];
    const int64_t n = (int64_t)fg.n[0] * fg.n[1] * fg.n[2];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int ix, iy, iz;
    dec3(t, fg.n[0], fg.n[1], ix, iy, iz);
    const float v = interp_f(d->g[coarse_level], cT, coordf(fg, 0, ix), coordf(fg, 1, iy),
                             coordf(fg, 2, iz));
    Tp[t] = fT[t] - v;
}

void launch_tprime(const GmStepDesc* d, int fine_level, int coarse_level, const float* fineT,
                   const float* coarseT, float* Tp, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_tprime, dim3((n + 255) / 256), dim3(256), 0, s, d, fine_level,
                       coarse_level, fineT, coarseT, Tp);
}

// getNewTprime in one launch when the coarse overlap nodes coincide with fine
// nodes (node-aligned levels): the injected coarse values (cF:2086-2094) are
// the fine values at those nodes, so every fine node reads them directly and
// the prolongation of cF:2097 needs no second launch.  Threads past the fine
// node count write the injected values into the coarse field.
__device__ __forceinline__ void tp_axis(int i, int f0, int R, float invR, int blo, int bn, int nc,
                                        int nf, int& c0, int& c1, float& w, bool& in0, bool& in1,
                                        int& j0, int& j1) {
    // 24-bit products (full rate): indices and ratios stay far below 2^23
    const int rel = i - f0 + R;                           // >= 1: f0 < R
    const int q = (int)(((float)rel + 0.5f) * invR);      // floor(rel / R), rel small
    w = (float)(rel - __mul24(q, R)) * invR;
    c0 = blo + q - 1;                                     // coarse node at or below node i
    c0 = c0 < 0 ? 0 : (c0 > nc - 1 ? nc - 1 : c0);        // inside the mesh when aligned
    c1 = c0 + 1 < nc ? c0 + 1 : nc - 1;                   // weight w = 0 when clamped
    j0 = f0 + __mul24(c0 - blo, R);
    j1 = f0 + __mul24(c1 - blo, R);
    // injected: in the overlap box and on a fine node
    in0 = c0 - blo >= 0 && c0 - blo < bn && j0 >= 0 && j0 < nf;
    in1 = c1 - blo >= 0 && c1 - blo < bn && j1 >= 0 && j1 < nf;
}

// Geometry of a fine/coarse pair for T' and injection, the same for every
// lane: kept in scalar registers (in the batched launch the descriptor is
// read through a pointer loaded from memory, where the compiler cannot see
// that the values are uniform; buffer resources and scalar offsets are built
// from them).
struct TpGeom {
    int f0[3], fr[3];  // fine node of overlap-box node 0; fine nodes per coarse spacing
    int fn[3], cn[3];  // fine and coarse node counts
    int blo[3], bn[3]; // overlap box in coarse nodes
};

__device__ __forceinline__ TpGeom tp_geom(const GmStepDesc* __restrict__ d, int fine_level,
                                          int coarse_level) {
    const gm_grid& fg = d->g[fine_level];
    const gm_grid& cg = d->g[coarse_level];
    const GmBox& b = fine_level == 3 ? d->ov3 : d->ov2;
    const gm_grid& xyz = fine_level == 3 ? d->ov3_xyz : d->ov2_xyz;
    TpGeom g;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        // integer offset and ratio: one hardware reciprocal is exact enough to round
        const float ih = __builtin_amdgcn_rcpf(fg.h[a]);
        g.f0[a] = __builtin_amdgcn_readfirstlane((int)rintf((xyz.x0[a] - fg.x0[a]) * ih));
        g.fr[a] = __builtin_amdgcn_readfirstlane((int)rintf(cg.h[a] * ih));
        g.fn[a] = __builtin_amdgcn_readfirstlane(fg.n[a]);
        g.cn[a] = __builtin_amdgcn_readfirstlane(cg.n[a]);
        g.blo[a] = __builtin_amdgcn_readfirstlane(b.lo[a]);
        g.bn[a] = __builtin_amdgcn_readfirstlane(b.n[a]);
    }
    return g;
}

// T' = T_fine - I(T_coarse after injection) at the fine nodes of block blk
// of a plane (z uniform in the block): with the coarse nodes on every R-th
// fine node, the coarse element of a fine node and its trilinear weights
// follow from the fine index; corners inside the injected box take the fine
// value (cF:2060-2099).  Buffer loads with 32-bit byte offsets, the plane's
// part of each offset in a scalar register.
__device__ __forceinline__ void tprime_blk(const TpGeom& g, const float* __restrict__ fT,
                                           const float* __restrict__ cT, float* __restrict__ Tp,
                                           int blk) {
    const int fnx = g.fn[0], fnxy = g.fn[0] * g.fn[1];
    const int nqb = (fnxy + 255) / 256;
    const int iz = blk / nqb;
    const int q = (blk - iz * nqb) * 256 + threadIdx.x;
    if (q >= fnxy) return;
    const int iy = fnxy < (1 << 20) ? fdiv_small(q, __builtin_amdgcn_rcpf((float)fnx))
                                    : (int)((unsigned)q / (unsigned)fnx);
    const int ix = q - iy * fnx;
    int cx0, cx1, cy0, cy1, cz0, cz1, jx0, jx1, jy0, jy1, jz0, jz1;
    float wx, wy, wz;
    bool ix0, ix1, iy0, iy1, iz0, iz1;
    tp_axis(ix, g.f0[0], g.fr[0], __builtin_amdgcn_rcpf((float)g.fr[0]), g.blo[0], g.bn[0],
            g.cn[0], g.fn[0], cx0, cx1, wx, ix0, ix1, jx0, jx1);
    tp_axis(iy, g.f0[1], g.fr[1], __builtin_amdgcn_rcpf((float)g.fr[1]), g.blo[1], g.bn[1],
            g.cn[1], g.fn[1], cy0, cy1, wy, iy0, iy1, jy0, jy1);
    tp_axis(iz, g.f0[2], g.fr[2], __builtin_amdgcn_rcpf((float)g.fr[2]), g.blo[2], g.bn[2],
            g.cn[2], g.fn[2], cz0, cz1, wz, iz0, iz1, jz0, jz1);
    // the z axis is the block's: its results to scalar registers
    cz0 = __builtin_amdgcn_readfirstlane(cz0);
    cz1 = __builtin_amdgcn_readfirstlane(cz1);
    jz0 = __builtin_amdgcn_readfirstlane(jz0);
    jz1 = __builtin_amdgcn_readfirstlane(jz1);
    wz = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(wz)));
    iz0 = __builtin_amdgcn_readfirstlane((int)iz0) != 0;
    iz1 = __builtin_amdgcn_readfirstlane((int)iz1) != 0;
    const uint32_t nf4 = 4u * (uint32_t)fnxy * (uint32_t)g.fn[2];
    const auto rF = buf_rsrc(fT, nf4);
    const uint32_t zt = 4u * (uint32_t)(fnxy * iz);
    const uint32_t z0 = 4u * (uint32_t)(fnxy * jz0), z1 = 4u * (uint32_t)(fnxy * jz1);
    const uint32_t y0 = 4u * __umul24((unsigned)fnx, (unsigned)jy0);
    const uint32_t y1 = 4u * __umul24((unsigned)fnx, (unsigned)jy1);
    const uint32_t x0 = 4u * (uint32_t)jx0, x1 = 4u * (uint32_t)jx1;
    float v;
    if (ix0 && ix1 && iy0 && iy1 && iz0 && iz1) {
        // all corners injected (inside the window): fine values only
        const float a00 = (1.0f - wx) * buf_ld(rF, x0 + y0, z0) + wx * buf_ld(rF, x1 + y0, z0);
        const float a10 = (1.0f - wx) * buf_ld(rF, x0 + y1, z0) + wx * buf_ld(rF, x1 + y1, z0);
        const float a01 = (1.0f - wx) * buf_ld(rF, x0 + y0, z1) + wx * buf_ld(rF, x1 + y0, z1);
        const float a11 = (1.0f - wx) * buf_ld(rF, x0 + y1, z1) + wx * buf_ld(rF, x1 + y1, z1);
        v = (1.0f - wz) * ((1.0f - wy) * a00 + wy * a10) + wz * ((1.0f - wy) * a01 + wy * a11);
    } else {
        // corners outside the injected box read the coarse field (whose
        // clamped node indices are inside it); both loads are issued, the
        // fine one of such a corner at an offset past the end (zero)
        const int cnx = g.cn[0], cnxy = g.cn[0] * g.cn[1];
        const auto rC = buf_rsrc(cT, 4u * (uint32_t)cnxy * (uint32_t)g.cn[2]);
        const uint32_t cy0o = 4u * __umul24((unsigned)cnx, (unsigned)cy0);
        const uint32_t cy1o = 4u * __umul24((unsigned)cnx, (unsigned)cy1);
        const uint32_t cz0o = 4u * (uint32_t)(cnxy * cz0), cz1o = 4u * (uint32_t)(cnxy * cz1);
        v = 0.0f;
#pragma unroll
        for (int qd = 0; qd < 8; ++qd) {
            const int dx = qd & 1, dy = (qd >> 1) & 1, dz = qd >> 2;
            const bool inj = (dx ? ix1 : ix0) && (dy ? iy1 : iy0) && (dz ? iz1 : iz0);
            // an injected corner is a node of the fine field
            const float uf =
                buf_ld(rF, inj ? (dx ? x1 : x0) + (dy ? y1 : y0) + (dz ? z1 : z0) : GM_OOB);
            const float uc = buf_ld(rC, 4u * (uint32_t)(dx ? cx1 : cx0) + (dy ? cy1o : cy0o),
                                    dz ? cz1o : cz0o);
            const float u = inj ? uf : uc;
            const float N =
                (dx ? wx : 1.0f - wx) * (dy ? wy : 1.0f - wy) * (dz ? wz : 1.0f - wz);
            v += N * u;
        }
    }
    buf_st(buf_ld(rF, 4u * (uint32_t)q, zt) - v, buf_rsrc(Tp, nf4), 4u * (uint32_t)q, zt);
}

// Injection of the fine values into the coarse overlap box, block ib of the
// box's nodes; the coarse level's carried k, cp follow the injected
// temperatures.
__device__ __forceinline__ void inject_blk(const TpGeom& g, const float* __restrict__ fT,
                                           float* __restrict__ cT, const DevMat* __restrict__ m,
                                           const uint8_t* __restrict__ cS1, int64_t c_nsub,
                                           float* __restrict__ cKC, int ib) {
    __shared__ float sTab[PTW] __attribute__((aligned(16)));
    if (cKC) {
        stage_tables(m, sTab, threadIdx.x, blockDim.x);
        __syncthreads();
    }
    const int k = ib * 256 + threadIdx.x;
    const int nb = g.bn[0] * g.bn[1] * g.bn[2];
    if (k >= nb) return;
    const int bz = k / (g.bn[0] * g.bn[1]), rem = k - bz * (g.bn[0] * g.bn[1]);
    const int by = rem / g.bn[0], bx = rem - by * g.bn[0];
    const int cx = g.blo[0] + bx, cy = g.blo[1] + by, cz = g.blo[2] + bz;
    if (cx < 0 || cx >= g.cn[0] || cy < 0 || cy >= g.cn[1] || cz < 0 || cz >= g.cn[2]) return;
    const int jx = g.f0[0] + bx * g.fr[0], jy = g.f0[1] + by * g.fr[1], jz = g.f0[2] + bz * g.fr[2];
    if (jx < 0 || jx >= g.fn[0] || jy < 0 || jy >= g.fn[1] || jz < 0 || jz >= g.fn[2]) return;
    const int64_t ci = nid(cx, cy, cz, g.cn[0], g.cn[1]);
    const float v = fT[jx + g.fn[0] * jy + g.fn[0] * g.fn[1] * jz];
    cT[ci] = v;
    if (cKC) {
        const int64_t nnc = (int64_t)g.cn[0] * g.cn[1] * g.cn[2];
        const PropC pc = load_pc(m);
        float kk, cc;
        bool s1;
        props_c(pc, sTab, v, cS1[ci] != 0, ci < c_nsub, kk, cc, s1);
        cKC[ci] = kk;
        cKC[nnc + ci] = cc;
    }
}

// Blocks below ntp_blocks: T' (one plane's 256 fine nodes each); the rest
// inject (256 overlap-box nodes each).
__global__ void k_tprime_inject(const GmStepDesc* __restrict__ d, int fine_level,
                                int coarse_level, const float* __restrict__ fT,
                                float* __restrict__ cT, float* __restrict__ Tp, int ntp_blocks,
                                const DevMat* __restrict__ m, const uint8_t* __restrict__ cS1,
                                int64_t c_nsub, float* __restrict__ cKC) {
    const TpGeom g = tp_geom(d, fine_level, coarse_level);
    const int blk = (int)blockIdx.x;
    if (blk < ntp_blocks) tprime_blk(g, fT, cT, Tp, blk);
    else inject_blk(g, fT, cT, m, cS1, c_nsub, cKC, blk - ntp_blocks);
}

// T' and injection of many engines' L3 windows into their L2 fields in one
// launch: bstart[i] is the first block of item i.  A block runs
// GM_TP_PER_BLOCK consecutive T' blocks of the single launch (the item
// lookup and the descriptor reads once for all of them, their loads in
// flight together), or one injection block.
#define GM_TP_PER_BLOCK 4
__global__ void __launch_bounds__(256) k_tprime_batch(const TpItem* __restrict__ it,
                                                      const int* __restrict__ bstart, int n) {
    const int b = (int)blockIdx.x;
    int b0;
    const int lo = batch_item_of(bstart, n, b, &b0);
    const TpItem x = it[lo];
    const TpGeom g = tp_geom(x.d, 3, 2);
    const int t = b - b0;
    const int ntpb = (x.ntp + GM_TP_PER_BLOCK - 1) / GM_TP_PER_BLOCK;
    if (t < ntpb) {
#pragma unroll
        for (int s = 0; s < GM_TP_PER_BLOCK; ++s) {
            const int blk = t * GM_TP_PER_BLOCK + s;
            if (blk < x.ntp) tprime_blk(g, x.fT, x.cT, x.Tp, blk);
        }
    } else {
        inject_blk(g, x.fT, x.cT, x.m, x.cS1, x.c_nsub, x.cKC, t - ntpb);
    }
}

void launch_tprime_batch(const TpItem* it, const int* bstart, int n, int nblocks,
                         hipStream_t s) {
    if (n <= 0 || nblocks <= 0) return;
    hipLaunchKernelGGL(k_tprime_batch, dim3(nblocks), dim3(256), 0, s, it, bstart, n);
}

int tprime_blocks(const gm_grid& fine_grid, int64_t nbox, int* ntp_out) {
    const int nxy = fine_grid.n[0] * fine_grid.n[1];
    const int ntp = ((nxy + 255) / 256) * fine_grid.n[2];
    if (ntp_out) *ntp_out = ntp;
    return (ntp + GM_TP_PER_BLOCK - 1) / GM_TP_PER_BLOCK + (int)((nbox + 255) / 256);
}

void launch_tprime_inject(const GmStepDesc* d, int fine_level, int coarse_level,
                          const float* fineT, float* coarseT, float* Tp, const gm_grid& fine_grid,
                          int64_t nbox, const DevMat* m, const uint8_t* cS1, int64_t c_nsub,
                          float* cKC, hipStream_t s) {
    // fine node counts are fixed for a level (moves change the origin only)
    const int nxy = fine_grid.n[0] * fine_grid.n[1];
    const int ntp = ((nxy + 255) / 256) * fine_grid.n[2];
    const int ninj = (int)((nbox + 255) / 256);
    hipLaunchKernelGGL(k_tprime_inject, dim3(ntp + ninj), dim3(256), 0, s, d, fine_level,
                       coarse_level, fineT, coarseT, Tp, ntp, m, cS1, c_nsub, cKC);
}

// moveEverything, L3 part (cF:2439-2444): T0 = I(L1 T0) + (I(L2 T'0) + I(L3 T'0))
// on the moved L3 nodes, with the pre-move L2 and L3 meshes.
__global__ void k_move_l3(const GmStepDesc* __restrict__ d, const float* __restrict__ T1, const float* __restrict__ Tp2,
                          const float* __restrict__ Tp3, float* __restrict__ T3n,
                          float* __restrict__ Tp3n, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const gm_grid& g3 = d->g[3];
    int ix, iy, iz;
    dec3(t, g3.n[0], g3.n[1], ix, iy, iz);
    const float x = coordf(g3, 0, ix), y = coordf(g3, 1, iy), z = coordf(g3, 2, iz);
    const float a = interp_f(d->g[1], T1, x, y, z);
    const float b = interp_f(d->gprev[2], Tp2, x, y, z);
    const float c = interp_f(d->gprev[3], Tp3, x, y, z);
    T3n[t] = a + (b + c);
    Tp3n[t] = c;
}

void launch_move_l3(const GmStepDesc* d, const float* T1, const float* Tp2, const float* Tp3,
                    float* T3n, float* Tp3n, int64_t nn3, hipStream_t s) {
    hipLaunchKernelGGL(k_move_l3, dim3((nn3 + 255) / 256), dim3(256), 0, s, d, T1, Tp2, Tp3, T3n,
                       Tp3n, nn3);
}

// moveEverything, L2 part (cF:2460-2464)
__global__ void k_move_l2(const GmStepDesc* __restrict__ d, const float* __restrict__ T1, const float* __restrict__ Tp2,
                          float* __restrict__ T2n, float* __restrict__ Tp2n, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const gm_grid& g2 = d->g[2];
    int ix, iy, iz;
    dec3(t, g2.n[0], g2.n[1], ix, iy, iz);
    const float x = coordf(g2, 0, ix), y = coordf(g2, 1, iy), z = coordf(g2, 2, iz);
    const float tp = interp_f(d->gprev[2], Tp2, x, y, z);
    const float a = interp_f(d->g[1], T1, x, y, z);
    Tp2n[t] = tp;
    T2n[t] = a + tp;
}

void launch_move_l2(const GmStepDesc* d, const float* T1, const float* Tp2, float* T2n,
                    float* Tp2n, int64_t nn2, hipStream_t s) {
    hipLaunchKernelGGL(k_move_l2, dim3((nn2 + 255) / 256), dim3(256), 0, s, d, T1, Tp2, T2n, Tp2n,
                       nn2);
}

// L1 S1 on the L2 overlap <- interpolated L2 S1 (cF:3273-3277)
__global__ void k_l1_s1_from_l2(const GmStepDesc* __restrict__ d, const uint8_t* __restrict__ S12,
                                float* __restrict__ S11) {
    const GmBox& b = d->ov2;
    const int n = b.n[0] * b.n[1] * b.n[2];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int ix, iy, iz;
    dec3(t, b.n[0], b.n[1], ix, iy, iz);
    const gm_grid& xyz = d->ov2_xyz;
    float N[8];
    int64_t nd[8];
    interp8(d->g[2], coordf(xyz, 0, ix), coordf(xyz, 1, iy), coordf(xyz, 2, iz), N, nd);
    float s = 0.0f;
#pragma unroll
    for (int a = 0; a < 8; ++a) s += N[a] * (float)S12[nd[a]];
    const gm_grid& g1 = d->g[1];
    const int cx = b.lo[0] + ix, cy = b.lo[1] + iy, cz = b.lo[2] + iz;
    if (cx < 0 || cx >= g1.n[0] || cy < 0 || cy >= g1.n[1] || cz < 0 || cz >= g1.n[2]) return;
    S11[nid(cx, cy, cz, g1.n[0], g1.n[1])] = s;
}

void launch_l1_s1_from_l2(const GmStepDesc* d, const uint8_t* S1_2, float* S1_1, int64_t n_box,
                          hipStream_t s) {
    if (n_box <= 0) return;
    hipLaunchKernelGGL(k_l1_s1_from_l2, dim3((n_box + 255) / 256), dim3(256), 0, s, d, S1_2, S1_1);
}

// ---------------------------------------------------------------------------
// Small utilities
// ---------------------------------------------------------------------------

__global__ void k_fill_f32(float* p, float v, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}
__global__ void k_fill_u8(uint8_t* p, uint8_t v, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}
__global__ void k_copy_f32(float* __restrict__ dst, const float* __restrict__ src, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = src[t];
}
__global__ void k_copy_u8(uint8_t* __restrict__ dst, const uint8_t* __restrict__ src, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = src[t];
}
__global__ void k_clamp(float* T, int64_t n, float Tamb) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n && T[t] < Tamb) T[t] = Tamb;
}

static inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

void launch_fill_f32(float* p, float v, int64_t n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_fill_f32, dim3(nblk(n)), dim3(256), 0, s, p, v, n);
}
void launch_fill_u8(uint8_t* p, uint8_t v, int64_t n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_fill_u8, dim3(nblk(n)), dim3(256), 0, s, p, v, n);
}
void launch_copy_f32(float* dst, const float* src, int64_t n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_copy_f32, dim3(nblk(n)), dim3(256), 0, s, dst, src, n);
}
void launch_copy_u8(uint8_t* dst, const uint8_t* src, int64_t n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_copy_u8, dim3(nblk(n)), dim3(256), 0, s, dst, src, n);
}
void launch_clamp(float* T, int64_t n, float Tamb, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_clamp, dim3(nblk(n)), dim3(256), 0, s, T, n, Tamb);
}

// Level0 <-> L3/L2 state transfer (cF:2500-2502, 2390-2392, 3628-3630).
// Level0 S2 is kept as a snapshot of the last L3 S2 and the Level0 box it was
// written to: the reference clears all of Level0 S2 before each write, so
// every Level0 node outside that box reads False.
__global__ void k_l0_gather(const GmStepDesc* __restrict__ d, const uint8_t* __restrict__ L0S1,
                            const uint8_t* __restrict__ snap, const int32_t* __restrict__ sbox,
                            uint8_t* __restrict__ S13, uint8_t* __restrict__ S23,
                            uint8_t* __restrict__ S12, int64_t nn3, int64_t nn2) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nn3) {
        const gm_grid& g3 = d->g[3];
        int ix, iy, iz;
        dec3(t, g3.n[0], g3.n[1], ix, iy, iz);
        const int lx = d->l0_lo3[0] + ix, ly = d->l0_lo3[1] + iy, lz = d->l0_lo3[2] + iz;
        const bool in0 = lx >= 0 && lx < d->l0_n[0] && ly >= 0 && ly < d->l0_n[1] && lz >= 0 &&
                         lz < d->l0_n[2];
        S13[t] = in0 ? L0S1[nid(lx, ly, lz, d->l0_n[0], d->l0_n[1])] : 0;
        uint8_t s2 = 0;
        if (sbox[3]) {
            const int sx = lx - sbox[0], sy = ly - sbox[1], sz = lz - sbox[2];
            if (sx >= 0 && sx < g3.n[0] && sy >= 0 && sy < g3.n[1] && sz >= 0 && sz < g3.n[2])
                s2 = snap[nid(sx, sy, sz, g3.n[0], g3.n[1])];
        }
        S23[t] = s2;
    }
    if (t < nn2) {
        const gm_grid& g2 = d->g[2];
        int ix, iy, iz;
        dec3(t, g2.n[0], g2.n[1], ix, iy, iz);
        const int lx = d->l0_lo2[0] + d->l0_st2[0] * ix, ly = d->l0_lo2[1] + d->l0_st2[1] * iy,
                  lz = d->l0_lo2[2] + d->l0_st2[2] * iz;
        const bool in0 = lx >= 0 && lx < d->l0_n[0] && ly >= 0 && ly < d->l0_n[1] && lz >= 0 &&
                         lz < d->l0_n[2];
        S12[t] = in0 ? L0S1[nid(lx, ly, lz, d->l0_n[0], d->l0_n[1])] : 0;
    }
}

void launch_l0_gather(const GmStepDesc* d, const uint8_t* L0S1, const uint8_t* s2snap,
                      const int32_t* s2box, uint8_t* S1_3, uint8_t* S2_3, uint8_t* S1_2,
                      int64_t nn3, int64_t nn2, hipStream_t s) {
    const int64_t n = nn3 > nn2 ? nn3 : nn2;
    hipLaunchKernelGGL(k_l0_gather, dim3(nblk(n)), dim3(256), 0, s, d, L0S1, s2snap, s2box, S1_3,
                       S2_3, S1_2, nn3, nn2);
}

__global__ void k_l0_scatter(const GmStepDesc* __restrict__ d, const uint8_t* __restrict__ S13,
                             const uint8_t* __restrict__ S23, uint8_t* __restrict__ L0S1,
                             uint8_t* __restrict__ snap, int32_t* __restrict__ sbox, int64_t nn3) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nn3) return;
    const gm_grid& g3 = d->g[3];
    int ix, iy, iz;
    dec3(t, g3.n[0], g3.n[1], ix, iy, iz);
    const int lx = d->l0_lo3[0] + ix, ly = d->l0_lo3[1] + iy, lz = d->l0_lo3[2] + iz;
    if (lx >= 0 && lx < d->l0_n[0] && ly >= 0 && ly < d->l0_n[1] && lz >= 0 && lz < d->l0_n[2])
        L0S1[nid(lx, ly, lz, d->l0_n[0], d->l0_n[1])] = S13[t];
    snap[t] = S23[t];
    if (t == 0) {
        sbox[0] = d->l0_lo3[0];
        sbox[1] = d->l0_lo3[1];
        sbox[2] = d->l0_lo3[2];
        sbox[3] = 1;
    }
}

void launch_l0_scatter(const GmStepDesc* d, const uint8_t* S1_3, const uint8_t* S2_3, uint8_t* L0S1,
                       uint8_t* s2snap, int32_t* s2box, int64_t nn3, hipStream_t s) {
    hipLaunchKernelGGL(k_l0_scatter, dim3(nblk(nn3)), dim3(256), 0, s, d, S1_3, S2_3, L0S1, s2snap,
                       s2box, nn3);
}

// Melt-time bookkeeping of the single-step driver path (go_melt.py:339-357).
__global__ void k_melt_single(const GmStepDesc* __restrict__ d, const float* __restrict__ T3,
                              const uint8_t* __restrict__ preS2, const uint8_t* __restrict__ S2,
                              float dt, float Tliq, float* __restrict__ accum,
                              float* __restrict__ maxacc, int64_t nn3) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nn3) return;
    const gm_grid& g3 = d->g[3];
    int ix, iy, iz;
    dec3(t, g3.n[0], g3.n[1], ix, iy, iz);
    const int lx = d->l0_lo3[0] + ix, ly = d->l0_lo3[1] + iy, lz = d->l0_lo3[2] + iz;
    if (lx < 0 || lx >= d->l0_n[0] || ly < 0 || ly >= d->l0_n[1] || lz < 0 || lz >= d->l0_n[2])
        return;
    const int64_t l0 = nid(lx, ly, lz, d->l0_n[0], d->l0_n[1]);
    const bool reset = (!preS2[t]) && S2[t];
    float acc = accum[l0];
    const float racc = reset ? acc : 0.0f;
    const float mx = maxacc[l0];
    maxacc[l0] = racc > mx ? racc : mx;
    acc = acc - racc;
    acc = acc + ((T3[t] > Tliq) ? 1.0f : 0.0f) * dt;
    accum[l0] = acc;
}

void launch_melt_single(const GmStepDesc* d, const float* T3, const uint8_t* preS2,
                        const uint8_t* S2, float dt, float Tliq, float* accum, float* max_accum,
                        int64_t nn3, hipStream_t s) {
    hipLaunchKernelGGL(k_melt_single, dim3(nblk(nn3)), dim3(256), 0, s, d, T3, preS2, S2, dt, Tliq,
                       accum, max_accum, nn3);
}

// min / max with NaN propagation (printLevelMaxMin, cF:3635-3665)
__global__ void k_minmax(const float* __restrict__ T, int64_t n, float* out) {
    __shared__ float smn[256], smx[256];
    float mn = INFINITY, mx = -INFINITY;
    bool nan = false;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        const float v = T[i];
        nan = nan || (v != v);
        mn = fminf(mn, v);
        mx = fmaxf(mx, v);
    }
    if (nan) {
        mn = NAN;
        mx = NAN;
    }
    smn[threadIdx.x] = mn;
    smx[threadIdx.x] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        float a = smn[0], b = smx[0];
        for (int i = 1; i < 256; ++i) {
            const float c = smn[i], e = smx[i];
            if (c != c || a != a) a = NAN; else a = fminf(a, c);
            if (e != e || b != b) b = NAN; else b = fmaxf(b, e);
        }
        out[0] = a;
        out[1] = b;
    }
}

void launch_minmax(const float* T, int64_t n, float* out2, hipStream_t s) {
    hipLaunchKernelGGL(k_minmax, dim3(1), dim3(256), 0, s, T, n, out2);
}

// ---- melt-pool size ----------------------------------------------------------
// The reference only forms S2 = T >= T_liquidus (cF:2585); the size below is
// this library's own measure.  Depth: the largest distance from the mesh's top
// plane down to where T crosses T_liquidus on a vertical edge whose upper node
// is molten (linear along the edge; a molten bottom node counts its full
// depth).  Width (length): the y (x) extent of the molten part of the top
// plane, its ends placed where T crosses T_liquidus along y (x) edges, or on
// the mesh boundary where a boundary node is molten.  Reductions in order-
// preserving integer keys of the FP32 values (one atomic per wave and value).
__device__ __forceinline__ int fkey(float f) {
    const int i = __float_as_int(f);
    return i >= 0 ? i : (i ^ 0x7fffffff);
}
__device__ __forceinline__ float kfloat(int k) { return __int_as_float(k >= 0 ? k : (k ^ 0x7fffffff)); }

// keys: [0] depth (max), [1] y lo (min), [2] y hi (max), [3] x lo (min),
// [4] x hi (max), [5] molten node count
__global__ void k_pool_init(int* keys) {
    // maxima start at INT_MIN, minima (odd slots 1, 3) at INT_MAX, the count at 0
    if (threadIdx.x < 6)
        keys[threadIdx.x] = threadIdx.x == 5 ? 0 : ((threadIdx.x & 1) ? INT_MAX : INT_MIN);
}

__global__ void __launch_bounds__(256) k_pool(gm_grid g, const float* __restrict__ T, float Tl,
                                              int* keys) {
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const int nxy = nx * ny;
    const int64_t nn = (int64_t)nxy * nz;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int kd = INT_MIN, kylo = INT_MAX, kyhi = INT_MIN, kxlo = INT_MAX, kxhi = INT_MIN, cnt = 0;
    if (t < nn) {
        int ix, iy, iz;
        dec3(t, nx, ny, ix, iy, iz);
        const float Ti = T[t];
        const bool mi = Ti >= Tl;
        cnt = mi ? 1 : 0;
        const float ztop = g.x0[2] + (float)(nz - 1) * g.h[2];
        if (mi) {
            if (iz == 0) {
                kd = fkey(ztop - g.x0[2]);
            } else {
                const float Tb = T[t - nxy];
                if (!(Tb >= Tl)) {
                    const float zc = coordf(g, 2, iz - 1) + g.h[2] * ((Tl - Tb) / (Ti - Tb));
                    kd = fkey(ztop - zc);
                }
            }
        }
        if (iz == nz - 1) {
            float ylo = INFINITY, yhi = -INFINITY, xlo = INFINITY, xhi = -INFINITY;
            if (mi && iy == 0) ylo = coordf(g, 1, 0);
            if (mi && iy == ny - 1) yhi = coordf(g, 1, ny - 1);
            if (mi && ix == 0) xlo = coordf(g, 0, 0);
            if (mi && ix == nx - 1) xhi = coordf(g, 0, nx - 1);
            if (iy + 1 < ny) {
                const float Tj = T[t + nx];
                if (mi != (Tj >= Tl)) {
                    const float yc = coordf(g, 1, iy) + g.h[1] * ((Tl - Ti) / (Tj - Ti));
                    ylo = fminf(ylo, yc);
                    yhi = fmaxf(yhi, yc);
                }
            }
            if (ix + 1 < nx) {
                const float Tj = T[t + 1];
                if (mi != (Tj >= Tl)) {
                    const float xc = coordf(g, 0, ix) + g.h[0] * ((Tl - Ti) / (Tj - Ti));
                    xlo = fminf(xlo, xc);
                    xhi = fmaxf(xhi, xc);
                }
            }
            if (ylo < INFINITY) kylo = fkey(ylo);
            if (yhi > -INFINITY) kyhi = fkey(yhi);
            if (xlo < INFINITY) kxlo = fkey(xlo);
            if (xhi > -INFINITY) kxhi = fkey(xhi);
        }
    }
#pragma unroll
    for (int sh = 32; sh >= 1; sh >>= 1) {
        kd = max(kd, __shfl_xor(kd, sh, 64));
        kylo = min(kylo, __shfl_xor(kylo, sh, 64));
        kyhi = max(kyhi, __shfl_xor(kyhi, sh, 64));
        kxlo = min(kxlo, __shfl_xor(kxlo, sh, 64));
        kxhi = max(kxhi, __shfl_xor(kxhi, sh, 64));
        cnt += __shfl_xor(cnt, sh, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (kd != INT_MIN) atomicMax(keys + 0, kd);
        if (kylo != INT_MAX) atomicMin(keys + 1, kylo);
        if (kyhi != INT_MIN) atomicMax(keys + 2, kyhi);
        if (kxlo != INT_MAX) atomicMin(keys + 3, kxlo);
        if (kxhi != INT_MIN) atomicMax(keys + 4, kxhi);
        if (cnt) atomicAdd(keys + 5, cnt);
    }
}

// keys -> dims: depth, width, length, molten node count, y lo, y hi, x lo, x hi
// (sizes 0 and ends NaN where nothing is molten)
__global__ void k_pool_finish(int* keys) {
    if (threadIdx.x != 0) return;
    const int k[6] = {keys[0], keys[1], keys[2], keys[3], keys[4], keys[5]};
    float* o = (float*)keys;
    const bool yok = k[1] != INT_MAX && k[2] != INT_MIN;
    const bool xok = k[3] != INT_MAX && k[4] != INT_MIN;
    o[0] = k[0] != INT_MIN ? kfloat(k[0]) : 0.0f;
    o[1] = yok ? kfloat(k[2]) - kfloat(k[1]) : 0.0f;
    o[2] = xok ? kfloat(k[4]) - kfloat(k[3]) : 0.0f;
    o[3] = (float)k[5];
    o[4] = yok ? kfloat(k[1]) : NAN;
    o[5] = yok ? kfloat(k[2]) : NAN;
    o[6] = xok ? kfloat(k[3]) : NAN;
    o[7] = xok ? kfloat(k[4]) : NAN;
}

void launch_melt_pool(const gm_grid& g, const float* T, float Tl, float* dims, hipStream_t s) {
    const int64_t nn = (int64_t)g.n[0] * g.n[1] * g.n[2];
    int* keys = (int*)dims;
    hipLaunchKernelGGL(k_pool_init, dim3(1), dim3(64), 0, s, keys);
    if (nn > 0) hipLaunchKernelGGL(k_pool, dim3(nblk(nn)), dim3(256), 0, s, g, T, Tl, keys);
    hipLaunchKernelGGL(k_pool_finish, dim3(1), dim3(64), 0, s, keys);
}

// ---- energy balance of one explicit step ----------------------------------------
// Audit of an explicit update T -> Tn of one level: with the lumped capacity
// C_i = sum_e rho*cp_e V_e/8 and the stiffness K of the element-mean k
// (the operator of solveMatrixFreeFE, cF:582-642), the unconstrained update is
// Tt = T + dt (F - K T) / C.  Because K annihilates constants (sum_i (K T)_i
// = 0), sum_i C_i (Tt - T)_i = dt sum_i F_i.  Nodes held by Dirichlet faces
// or by the T >= T_clamp floor are counted with Tt (their reaction is
// reported apart); every other node with Tn.  The relative residual
// (sum - dt sum F) / (sum of absolute energy flows) is FP32 rounding for the
// plain solver and measures the energy error of any other update (e.g. a
// surrogate).  Per-node terms and sums in FP64, one atomic per wave and sum.
// sums: [0] sum_free C (Tn - T), [1] sum_held C (Tt - T), [2] dt sum F,
// [3] dt sum K T, [4] sum_held C (Tn - Tt), [5] scale, [6] residual,
// [7] held nodes.
__global__ void k_energy_init(double* sums) {
    if (threadIdx.x < 8) sums[threadIdx.x] = 0.0;
}

__global__ void __launch_bounds__(256) k_energy(gm_grid g, int64_t ne_active, EnergyCoef ec,
                                                const float* __restrict__ k,
                                                const float* __restrict__ rc,
                                                const float* __restrict__ T,
                                                const float* __restrict__ F,
                                                const float* __restrict__ Corr,
                                                const float* __restrict__ Tn, int bc_faces,
                                                float T_clamp, double dt, double* sums) {
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const int nex = nx - 1, ney = ny - 1, nez = nz - 1;
    const int64_t nn = (int64_t)nx * ny * nz;
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nn; t += stride) {
        int ix, iy, iz;
        dec3(t, nx, ny, ix, iy, iz);
        double KT = 0.0, C = 0.0;
        for (int e = 0; e < 8; ++e) {
            // element with min corner (ix - 1 + ax, ...); this node is its local corner (1-ax, ...)
            const int ax = e & 1, ay = (e >> 1) & 1, az = e >> 2;
            const int ex = ix - 1 + ax, ey = iy - 1 + ay, ez = iz - 1 + az;
            if (ex < 0 || ex >= nex || ey < 0 || ey >= ney || ez < 0 || ez >= nez) continue;
            if ((int64_t)ex + (int64_t)nex * ((int64_t)ey + (int64_t)ney * ez) >= ne_active) continue;
            double ks = 0.0, cs = 0.0, kt = 0.0;
            for (int m = 0; m < 8; ++m) {
                const int mx = m & 1, my = (m >> 1) & 1, mz = m >> 2;
                const int64_t j = nid(ex + mx, ey + my, ez + mz, nx, ny);
                ks += (double)k[j];
                cs += (double)rc[j];
                const int pat = (mx ^ (1 - ax)) | ((my ^ (1 - ay)) << 1) | ((mz ^ (1 - az)) << 2);
                kt += ec.kc[pat] * (double)T[j];
            }
            KT += (ks * 0.125) * kt;
            C += (cs * 0.125) * ec.vol8;
        }
        if (C > 0.0) {
            const double Ti = T[t], Tni = Tn[t];
            const double Fi = (F ? (double)F[t] : 0.0) + (Corr ? (double)Corr[t] : 0.0);
            const double Tt = Ti + dt * (Fi - KT) / C;
            const bool face = ((bc_faces & 1) && ix == 0) || ((bc_faces & 2) && ix == nex) ||
                              ((bc_faces & 4) && iy == 0) || ((bc_faces & 8) && iy == ney) ||
                              ((bc_faces & 16) && iz == 0);
            const bool held = face || (Tni <= (double)T_clamp && Tt < (double)T_clamp);
            const double dE = held ? C * (Tt - Ti) : C * (Tni - Ti);
            if (held) {
                v[1] += dE;
                v[4] += C * (Tni - Tt);
                v[7] += 1.0;
            } else {
                v[0] += dE;
            }
            v[2] += dt * Fi;
            v[3] += dt * KT;
            v[5] += fabs(dt * Fi) + fabs(dE);
        }
    }
    // the eight sums share one cache line, so atomics on them serialise:
    // fold the block's four waves in LDS and add once per block
    __shared__ double part[4][8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        double x = v[q];
#pragma unroll
        for (int sh = 32; sh >= 1; sh >>= 1) x += __shfl_xor(x, sh, 64);
        if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6][q] = x;
    }
    __syncthreads();
    if (threadIdx.x < 8 && threadIdx.x != 6) {
        const int q = threadIdx.x;
        const double x = part[0][q] + part[1][q] + part[2][q] + part[3][q];
        if (x != 0.0) atomicAdd(sums + q, x);
    }
}

__global__ void k_energy_finish(double* sums) {
    if (threadIdx.x != 0) return;
    const double sc = sums[5];
    sums[6] = sc > 0.0 ? (sums[0] + sums[1] - sums[2]) / sc : 0.0;
}

EnergyCoef energy_coef(const gm_grid& g) {
    EnergyCoef c;
    const double h[3] = {g.h[0], g.h[1], g.h[2]};
    for (int p = 0; p < 8; ++p) {
        // box element, unit k: K = K1 (x) M1 (x) M1 + M1 (x) K1 (x) M1 + M1 (x) M1 (x) K1
        // with K1 = [[1,-1],[-1,1]]/h and M1 = h [[2,1],[1,2]]/6, entry by neighbour pattern p
        double kk[3], mm[3];
        for (int d = 0; d < 3; ++d) {
            const bool dif = (p >> d) & 1;
            kk[d] = (dif ? -1.0 : 1.0) / h[d];
            mm[d] = h[d] * (dif ? 1.0 : 2.0) / 6.0;
        }
        c.kc[p] = kk[0] * mm[1] * mm[2] + mm[0] * kk[1] * mm[2] + mm[0] * mm[1] * kk[2];
    }
    c.vol8 = h[0] * h[1] * h[2] / 8.0;
    return c;
}

void launch_energy(const gm_grid& g, int64_t ne_active, const float* k, const float* rc,
                   const float* T, const float* F, const float* Corr, const float* Tn,
                   int bc_faces, float T_clamp, float dt, double* sums, hipStream_t s) {
    const int64_t nn = (int64_t)g.n[0] * g.n[1] * g.n[2];
    hipLaunchKernelGGL(k_energy_init, dim3(1), dim3(64), 0, s, sums);
    if (nn > 0)
        hipLaunchKernelGGL(k_energy, dim3((unsigned)std::min<int64_t>(nblk(nn), 512)), dim3(256), 0,
                           s, g, ne_active, energy_coef(g),
                           k, rc, T, F, Corr, Tn, bc_faces, T_clamp, (double)dt, sums);
    hipLaunchKernelGGL(k_energy_finish, dim3(1), dim3(64), 0, s, sums);
}

// ---- single-operator kernels (the facade's one-function entry points) ----

__global__ void k_props(DevMat m, const float* __restrict__ T, const float* __restrict__ S1,
                        int64_t n, int64_t nsub, float* S1o, uint8_t* S2o, float* k, float* c) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float kv, cv;
    bool s1, s2;
    props(m, T[t], S1[t], t < nsub, kv, cv, s1, s2);
    if (S1o) S1o[t] = s1 ? 1.0f : 0.0f;
    if (S2o) S2o[t] = (uint8_t)s2;
    if (k) k[t] = kv;
    if (c) c[t] = cv;
}

void launch_props(const DevMat& m, const float* T, const float* S1, int64_t nn, int64_t nsub,
                  float* S1o, uint8_t* S2o, float* k, float* c, hipStream_t s) {
    hipLaunchKernelGGL(k_props, dim3(nblk(nn)), dim3(256), 0, s, m, T, S1, nn, nsub, S1o, S2o, k, c);
}

__global__ void k_source_l3_full(gm_grid g, DevMat m, float vx, float vy, float vz, float P,
                                 float w3, float* __restrict__ F) {
    const int64_t n = (int64_t)g.n[0] * g.n[1] * g.n[2];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int ix, iy, iz;
    dec3(t, g.n[0], g.n[1], ix, iy, iz);
    const float sx = src_axis(g, 0, ix, vx, m.rsq, m.rc) * (m.pc6 * P);
    const float sy = src_axis(g, 1, iy, vy, m.rsq, m.rc);
    const float sz = src_axis(g, 2, iz, vz, m.dsq, m.dc);
    F[t] = sx * sy * sz * w3;
}

void launch_source_l3_full(const gm_grid& g, const DevMat& m, float vx, float vy, float vz, float P,
                           float* F, hipStream_t s) {
    const int64_t n = (int64_t)g.n[0] * g.n[1] * g.n[2];
    const float w3 = (g.h[0] * 0.5f) * (g.h[1] * 0.5f) * (g.h[2] * 0.5f);
    hipLaunchKernelGGL(k_source_l3_full, dim3(nblk(n)), dim3(256), 0, s, g, m, vx, vy, vz, P, w3, F);
}

// Source averaged over up to 32 laser positions (one launch per pack):
// F += a * source(position r) in position order, F cleared by the first pack.
struct SrcPack {
    float x[32], y[32], z[32], P[32];
    int n;
    float a;
    int accumulate;
};

__global__ void k_source_l3_mean(gm_grid g, DevMat m, SrcPack pk, float w3,
                                 float* __restrict__ F) {
    const int64_t n = (int64_t)g.n[0] * g.n[1] * g.n[2];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int ix, iy, iz;
    dec3(t, g.n[0], g.n[1], ix, iy, iz);
    float acc = pk.accumulate ? F[t] : 0.0f;
    for (int r = 0; r < pk.n; ++r) {
        const float sx = src_axis(g, 0, ix, pk.x[r], m.rsq, m.rc) * (m.pc6 * pk.P[r]);
        const float sy = src_axis(g, 1, iy, pk.y[r], m.rsq, m.rc);
        const float sz = src_axis(g, 2, iz, pk.z[r], m.dsq, m.dc);
        acc += pk.a * (sx * sy * sz * w3);
    }
    F[t] = acc;
}

void launch_source_l3_mean(const gm_grid& g, const DevMat& m, const float* rows, int n_rows,
                           float* F, hipStream_t s) {
    const int64_t n = (int64_t)g.n[0] * g.n[1] * g.n[2];
    const float w3 = (g.h[0] * 0.5f) * (g.h[1] * 0.5f) * (g.h[2] * 0.5f);
    SrcPack pk;
    pk.a = 1.0f / (float)n_rows;
    for (int r0 = 0; r0 < n_rows; r0 += 32) {
        pk.n = n_rows - r0 < 32 ? n_rows - r0 : 32;
        pk.accumulate = r0 > 0;
        for (int r = 0; r < pk.n; ++r) {
            const float* row = rows + 7 * (r0 + r);
            pk.x[r] = row[0];
            pk.y[r] = row[1];
            pk.z[r] = row[2];
            pk.P[r] = row[6];
        }
        hipLaunchKernelGGL(k_source_l3_mean, dim3(nblk(n)), dim3(256), 0, s, g, m, pk, w3, F);
    }
}

__global__ void k_convrad_full(gm_grid g, DevMat m, const float* __restrict__ T, int top,
                               float w2, float* __restrict__ F) {
    const int nx = g.n[0], ny = g.n[1];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nx * ny) return;
    const int ix = t % nx, iy = t / nx;
    float cr = 0.0f;
    for (int e = 0; e < 4; ++e) {
        const int ex = ix - 1 + (e & 1), ey = iy - 1 + (e >> 1);
        if (ex < 0 || ex >= nx - 1 || ey < 0 || ey >= ny - 1) continue;
        const float Tf[4] = {T[nid(ex, ey, top, nx, ny)], T[nid(ex + 1, ey, top, nx, ny)],
                             T[nid(ex + 1, ey + 1, top, nx, ny)], T[nid(ex, ey + 1, top, nx, ny)]};
        const int ax = ix - ex, ay = iy - ey;
        const int ai = ay ? (ax ? 2 : 3) : (ax ? 1 : 0);
        for (int q = 0; q < 4; ++q) {
            const float Tq =
                n2(q, 0) * Tf[0] + n2(q, 1) * Tf[1] + n2(q, 2) * Tf[2] + n2(q, 3) * Tf[3];
            cr += n2(q, ai) * (face_flux(m, Tq) * w2);
        }
    }
    F[nid(ix, iy, top, nx, ny)] += cr;
}

void launch_convrad_full(const gm_grid& g, const DevMat& m, const float* T, int64_t ne_active,
                         float* F, hipStream_t s) {
    const int top = (int)(ne_active / ((int64_t)(g.n[0] - 1) * (g.n[1] - 1)));
    const float w2 = (g.h[0] * 0.5f) * (g.h[1] * 0.5f);
    const int n = g.n[0] * g.n[1];
    hipLaunchKernelGGL(k_convrad_full, dim3(nblk(n)), dim3(256), 0, s, g, m, T, top, w2, F);
}

__global__ void k_interp_matrix(gm_grid src, gm_grid tg, float* __restrict__ Nc,
                                int32_t* __restrict__ node) {
    const int64_t n = (int64_t)tg.n[0] * tg.n[1] * tg.n[2];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int ix, iy, iz;
    dec3(t, tg.n[0], tg.n[1], ix, iy, iz);
    float N[8];
    int64_t nd[8];
    interp8(src, coordf(tg, 0, ix), coordf(tg, 1, iy), coordf(tg, 2, iz), N, nd);
    for (int a = 0; a < 8; ++a) {
        Nc[t * 8 + a] = N[a];
        node[t * 8 + a] = (int32_t)nd[a];
    }
}

void launch_interp_matrix(const gm_grid& src, const gm_grid& tgt, float* Nc, int32_t* node,
                          hipStream_t s) {
    const int64_t n = (int64_t)tgt.n[0] * tgt.n[1] * tgt.n[2];
    hipLaunchKernelGGL(k_interp_matrix, dim3(nblk(n)), dim3(256), 0, s, src, tgt, Nc, node);
}

__global__ void k_interp_w(const float* __restrict__ Nc, const int32_t* __restrict__ node,
                           int64_t n, const float* __restrict__ T, float* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float s = 0.0f;
    for (int a = 0; a < 8; ++a) s += Nc[t * 8 + a] * T[node[t * 8 + a]];
    out[t] = s;
}

void launch_interp_w(const float* Nc, const int32_t* node, int64_t n, const float* T, float* out,
                     hipStream_t s) {
    hipLaunchKernelGGL(k_interp_w, dim3(nblk(n)), dim3(256), 0, s, Nc, node, n, T, out);
}

// Level0 layer shift (go_melt.py:264-290): v[:n2] = v[n1:]; v[n2:] = 0.
// Run as two passes through a scratch copy held in the upper part.
__global__ void k_shift_down_u8(uint8_t* __restrict__ dst, const uint8_t* __restrict__ src,
                                int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = src[t];
}

void launch_shift_l0(uint8_t* L0S1, float* accum, int64_t n1, int64_t n2, int64_t nn,
                     hipStream_t s) {
    // n1 <= n2 and the copy goes downwards, so chunks of n1 nodes copied in
    // order never overwrite a source chunk that is still to be read.
    for (int64_t off = 0; off < n2; off += n1) {
        const int64_t len = (off + n1 <= n2) ? n1 : (n2 - off);
        hipLaunchKernelGGL(k_shift_down_u8, dim3(nblk(len)), dim3(256), 0, s, L0S1 + off,
                           L0S1 + off + n1, len);
        if (accum)
            hipLaunchKernelGGL(k_copy_f32, dim3(nblk(len)), dim3(256), 0, s, accum + off,
                               accum + off + n1, len);
    }
    launch_fill_u8(L0S1 + n2, 0, nn - n2, s);
    if (accum) launch_fill_f32(accum + n2, 0.0f, nn - n2, s);
}

// ---------------------------------------------------------------------------
// Host-side constants
// ---------------------------------------------------------------------------

DevMat make_devmat(const gm_material& mm) {
    DevMat m{};
    m.T_amb = mm.T_amb;
    m.T_sol = mm.T_solidus;
    m.T_liq = mm.T_liquidus;
    m.T_boil = mm.T_boiling;
    m.k_p = mm.k_powder;
    m.k_a0 = mm.k_bulk_a0;
    m.k_a1 = mm.k_bulk_a1;
    m.k_f = mm.k_fluid;
    m.cp_a0 = mm.cp_solid_a0;
    m.cp_a1 = mm.cp_solid_a1;
    m.cp_m = mm.cp_mushy;
    m.cp_f = mm.cp_fluid;
    m.rho = mm.rho;
    m.k_n = mm.k_table_n;
    m.cp_n = mm.cp_table_n;
    for (int i = 0; i < GM_MAX_KNOTS; ++i) {
        m.kT[i] = mm.k_table_T[i];
        m.kv[i] = mm.k_table_v[i];
        m.cT[i] = mm.cp_table_T[i];
        m.cv[i] = mm.cp_table_v[i];
        m.ks[i] = 0.0f;
        m.cs[i] = 0.0f;
    }
    for (int i = 0; i + 1 < mm.k_table_n; ++i)
        m.ks[i] = (float)(((double)mm.k_table_v[i + 1] - mm.k_table_v[i]) /
                          ((double)mm.k_table_T[i + 1] - mm.k_table_T[i]));
    for (int i = 0; i + 1 < mm.cp_table_n; ++i)
        m.cs[i] = (float)(((double)mm.cp_table_v[i + 1] - mm.cp_table_v[i]) /
                          ((double)mm.cp_table_T[i + 1] - mm.cp_table_T[i]));
    m.h_conv = mm.h_conv;
    m.sigma_eps = mm.sigma_eps;
    m.evc_cp = mm.evc_cp;
    m.CT = mm.CT;
    m.CM = mm.CM;
    m.Lev = mm.Lev;
    m.invTb = (float)(1.0 / (double)mm.T_boiling);
    m.Tclip = (float)((double)mm.T_boiling + 1000.0);
    m.Tamb4 = (float)((double)mm.T_amb * mm.T_amb * mm.T_amb * mm.T_amb);
    m.sqCM = (float)sqrt((double)mm.CM);
    m.eta = mm.laser_eta;
    const double r = mm.laser_radius, dd = mm.laser_depth;
    m.rsq = (float)(r * r);
    m.dsq = (float)(dd * dd);
    m.rc = (float)(1.0 / (r * sqrt(M_PI)));
    m.dc = (float)(1.0 / (dd * sqrt(M_PI)));
    m.pc6 = (float)(6.0 * sqrt(3.0) * mm.laser_eta);
    // Union of the table knots; k and cp are linear on every segment between
    // consecutive knots (a property without a table keeps its linear law).
    std::vector<double> U;
    for (int i = 0; i < mm.k_table_n && mm.k_table_n >= 2; ++i) U.push_back(mm.k_table_T[i]);
    for (int i = 0; i < mm.cp_table_n && mm.cp_table_n >= 2; ++i) U.push_back(mm.cp_table_T[i]);
    std::sort(U.begin(), U.end());
    U.erase(std::unique(U.begin(), U.end()), U.end());
    if (U.size() > 2 * GM_MAX_KNOTS) U.resize(2 * GM_MAX_KNOTS);
    const int nU = (int)U.size();
    // value and slope on the segment that starts at u: table segment j of the
    // largest interior knot <= u (cF:2584-2612 with a table for the law)
    auto seg = [](const float* kT, const float* kv, int n, double a0, double a1, double u,
                  double& v, double& sl) {
        if (n < 2) {
            sl = a1;
            v = a0 + a1 * u;
            return;
        }
        int j = 0;
        for (int i = 1; i < n - 1; ++i)
            if ((double)kT[i] <= u) j = i;
        sl = ((double)kv[j + 1] - kv[j]) / ((double)kT[j + 1] - kT[j]);
        v = (double)kv[j] + sl * (u - kT[j]);
    };
    // segment of T on the union knots: largest i in [0, nU-2] with U_i <= T
    auto useg = [&](double T) {
        int j = 0;
        for (int i = 1; i < nU - 1; ++i)
            if (U[i] <= T) j = i;
        return j;
    };
    double segv[2 * GM_MAX_KNOTS][5];
    for (int j = 0; j < nU; ++j) {
        segv[j][0] = U[j];
        seg(mm.k_table_T, mm.k_table_v, mm.k_table_n, mm.k_bulk_a0, mm.k_bulk_a1, U[j],
            segv[j][1], segv[j][2]);
        seg(mm.cp_table_T, mm.cp_table_v, mm.cp_table_n, mm.cp_solid_a0, mm.cp_solid_a1, U[j],
            segv[j][3], segv[j][4]);
        for (int f = 0; f < 5; ++f) m.mseg[j][f] = (float)segv[j][f];
    }
    m.mn = nU >= 2 ? nU : 0;
    m.pnb = 0;
    if (nU >= 2) {
        const double range = U[nU - 1] - U[0];
        double dmin = range;
        for (int i = 0; i + 1 < nU; ++i) dmin = std::min(dmin, U[i + 1] - U[i]);
        // loop lookup: lut[b] is the segment half a bin below the bin's lower
        // edge, a lower bound of the segment of every T in bin b
        int nb = (int)std::min<double>(GM_LUT_BINS, std::ceil(2.0 * range / dmin));
        nb = std::max(nb, 1);
        double iw = nb / range;
        m.mT0 = (float)U[0];
        m.miw = (float)iw;
        m.mnb = nb;
        for (int b = 0; b < GM_LUT_BINS; ++b) m.mlut[b] = useg(U[0] + (b - 0.5) / iw);
        // records: the fewest bins for which every widened bin meets at most
        // two segments (also for T beyond the end knots in the end bins)
        for (int pnb = std::max(1, (int)std::ceil(range / dmin)); pnb <= GM_PREC_BINS; ++pnb) {
            const double w = range / pnb, dl = 0.25 * w;
            bool ok = true;
            int A[GM_PREC_BINS];
            for (int b = 0; b < pnb && ok; ++b) {
                const double lo = U[0] + b * w, hi = lo + w;
                A[b] = useg(lo - dl);
                const int top = (b == pnb - 1) ? nU - 2 : useg(hi + dl);
                ok = top <= A[b] + 1;
            }
            if (!ok) continue;
            m.pnb = pnb;
            m.pT0 = (float)U[0];
            m.piw = (float)(pnb / range);
            for (int b = 0; b < GM_PREC_BINS; ++b) {
                const int a = A[std::min(b, pnb - 1)];
                const int c = std::min(a + 1, nU - 2);
                float* r = m.prec[b];
                r[0] = a + 1 <= nU - 2 ? (float)U[a + 1] : INFINITY;
                for (int f = 0; f < 5; ++f) r[1 + f] = (float)segv[a][f];
                for (int f = 0; f < 4; ++f) r[6 + f] = (float)segv[c][1 + f];
                for (int f = 10; f < 16; ++f) r[f] = 0.0f;
            }
            break;
        }
    } else {
        // no table: one record of the two linear laws
        m.pnb = 1;
        m.pT0 = 0.0f;
        m.piw = 0.0f;
        for (int b = 0; b < GM_PREC_BINS; ++b) {
            float* r = m.prec[b];
            r[0] = INFINITY;
            r[1] = 0.0f;
            r[2] = mm.k_bulk_a0;
            r[3] = mm.k_bulk_a1;
            r[4] = mm.cp_solid_a0;
            r[5] = mm.cp_solid_a1;
            for (int f = 6; f < 16; ++f) r[f] = 0.0f;
        }
    }
    return m;
}

// Box-element operator: K_ab = k * kc[p], p = which axes a and b differ on;
// exact for the 2x2x2 Gauss rule of cF:692-775 (tensor products of the 1-D
// linear element: stiffness (1/h)[1 -1; -1 1], mass (h/6)[2 1; 1 2]).
void level_coeffs(const gm_grid& g, float kc[8], float* mrow, float* w2, float* w3) {
    const double h[3] = {g.h[0], g.h[1], g.h[2]};
    for (int p = 0; p < 8; ++p) {
        double kk[3], mm[3];
        for (int d = 0; d < 3; ++d) {
            const bool dif = (p >> d) & 1;
            kk[d] = (dif ? -1.0 : 1.0) / h[d];
            mm[d] = h[d] * (dif ? 1.0 / 6.0 : 1.0 / 3.0);
        }
        kc[p] = (float)(kk[0] * mm[1] * mm[2] + mm[0] * kk[1] * mm[2] + mm[0] * mm[1] * kk[2]);
    }
    *mrow = (float)(h[0] * h[1] * h[2] / 8.0);
    *w2 = (float)((h[0] / 2) * (h[1] / 2));
    *w3 = (float)((h[0] / 2) * (h[1] / 2) * (h[2] / 2));
}

}  // namespace gm
