// CPU restatement of the reference's hot functions in C++/OpenMP, FP32.
//
// Test and measurement infrastructure: the oracle's orchestration
// (oracle/gomelt_oracle.py) calls these through oracle/cpu_port.py to time the
// reference algorithm on all host cores as the CPU baseline of the benchmark.
// Nothing in the product path links it.
//
// Each function keeps the reference's structure: a loop over elements that
// gathers the element's nodal values, evaluates at the 2x2x2 Gauss points and
// scatters into nodal arrays (jnp.bincount / BCOO products of
// go_melt/computeFunctions.py, here cF).  Scatters are made race-free by
// colouring element rows (parity of y and z), or by atomic adds where the
// target is a coarser level.  Arithmetic is FP32, as the reference's.
//
// Meshes are passed as their three node-coordinate axes (x-fastest node
// numbering i = ix + nx*(iy + ny*iz), cF:645-689).

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include <omp.h>

namespace {

const double kXB[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
const double kYB[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
const double kZB[8] = {-1, -1, -1, -1, 1, 1, 1, 1};

struct Mesh {
    const float* ax[3];
    int n[3];    // nodes per axis
    int ne[3];   // elements per axis
    float h[3];  // element size from element 0 (cF:3135-3148)
    int64_t nn() const { return (int64_t)n[0] * n[1] * n[2]; }
};

Mesh mesh(const float* x, const float* y, const float* z, const int32_t* n) {
    Mesh m;
    m.ax[0] = x;
    m.ax[1] = y;
    m.ax[2] = z;
    for (int d = 0; d < 3; ++d) {
        m.n[d] = n[d];
        m.ne[d] = n[d] - 1;
        m.h[d] = m.ax[d][1] - m.ax[d][0];
    }
    return m;
}

// Reference element at 2-point Gauss (cF:692-775): N[q][a], dN[q][a][d], w.
// The Jacobian is the diagonal one of element 0.
struct RefElem {
    float N[8][8];
    float dN[8][8][3];
    float w;
};

RefElem ref_element(const Mesh& m) {
    RefElem r;
    const double g = 1.0 / sqrt(3.0);
    const double j[3] = {0.5 * m.h[0], 0.5 * m.h[1], 0.5 * m.h[2]};
    for (int q = 0; q < 8; ++q) {
        const double k = g * kXB[q], e = g * kYB[q], z = g * kZB[q];
        for (int a = 0; a < 8; ++a) {
            const double fk = 1 + k * kXB[a], fe = 1 + e * kYB[a], fz = 1 + z * kZB[a];
            r.N[q][a] = (float)(0.125 * fk * fe * fz);
            r.dN[q][a][0] = (float)(0.125 * kXB[a] * fe * fz / j[0]);
            r.dN[q][a][1] = (float)(0.125 * kYB[a] * fk * fz / j[1]);
            r.dN[q][a][2] = (float)(0.125 * kZB[a] * fk * fe / j[2]);
        }
    }
    r.w = (float)(j[0] * j[1] * j[2]);
    return r;
}

// Node ids of element (ix, iy, iz) in the local order of cF:645-689.
inline void elem_nodes(const Mesh& m, int ix, int iy, int iz, int64_t id[8]) {
    const int64_t nx = m.n[0], nxy = (int64_t)m.n[0] * m.n[1];
    const int64_t b = ix + nx * iy + nxy * iz;
    id[0] = b;
    id[1] = b + 1;
    id[2] = b + 1 + nx;
    id[3] = b + nx;
    id[4] = b + nxy;
    id[5] = b + 1 + nxy;
    id[6] = b + 1 + nx + nxy;
    id[7] = b + nx + nxy;
}

// Element rows (iy, iz) of the first `ne` elements, in 4 colours: rows of one
// colour share no node, so a thread owning a row scatters without races.
template <class F>
void for_rows_coloured(const Mesh& m, int64_t ne, F&& body) {
    const int64_t exy = (int64_t)m.ne[0] * m.ne[1];
    const int nzl = (int)std::min<int64_t>(m.ne[2], (ne + exy - 1) / exy);
    for (int c = 0; c < 4; ++c) {
        const int cy = c & 1, cz = c >> 1;
        const int ny2 = (m.ne[1] - cy + 1) / 2, nz2 = (nzl - cz + 1) / 2;
        if (ny2 <= 0 || nz2 <= 0) continue;
#pragma omp parallel for schedule(dynamic, 4)
        for (int64_t r = 0; r < (int64_t)ny2 * nz2; ++r) {
            const int iy = cy + 2 * (int)(r % ny2), iz = cz + 2 * (int)(r / ny2);
            const int64_t e0 = (int64_t)m.ne[0] * (iy + (int64_t)m.ne[1] * iz);
            if (e0 >= ne) continue;
            const int nx_e = (int)std::min<int64_t>(m.ne[0], ne - e0);
            body(iy, iz, nx_e);
        }
    }
}

inline void atomic_add(float* p, float v) {
#pragma omp atomic
    *p += v;
}

// Coarse element of a point along one axis, clipped (cF:1069-1077).
inline int coarse_elem(const Mesh& c, int d, float x) {
    const float h = c.h[d];
    int e = (int)floorf((x - c.ax[d][0]) / h);
    return std::min(std::max(e, 0), c.ne[d] - 1);
}

struct Src {
    float pc, rc, dc, rsq, dsq;
};

Src src_consts(const float* p) {
    // p: [eta, radius, depth]; computeSourceFunction_jax cF:991-1025
    Src s;
    s.pc = (float)(6.0 * sqrt(3.0)) * p[0];
    s.rc = (float)(1.0 / (p[1] * sqrt(M_PI)));
    s.dc = (float)(1.0 / (p[2] * sqrt(M_PI)));
    s.rsq = (float)((double)p[1] * p[1]);
    s.dsq = (float)((double)p[2] * p[2]);
    return s;
}

inline float source_q(const Src& s, float P, const float* v, float x, float y, float z) {
    const float dx = x - v[0], dy = y - v[1], dz = z - v[2];
    const float Qx = s.rc * expf(-3.0f * dx * dx / s.rsq);
    const float Qy = s.rc * expf(-3.0f * dy * dy / s.rsq);
    const float Qz = s.dc * expf(-3.0f * dz * dz / s.dsq);
    return (s.pc * P) * Qx * Qy * Qz;
}

// Gauss point coordinates of element (ix, iy, iz): x_q = sum_a N[q][a] x_a.
inline void gauss_xyz(const Mesh& m, const RefElem& r, int ix, int iy, int iz, float xq[8],
                      float yq[8], float zq[8]) {
    const float x0 = m.ax[0][ix], x1 = m.ax[0][ix + 1];
    const float y0 = m.ax[1][iy], y1 = m.ax[1][iy + 1];
    const float z0 = m.ax[2][iz], z1 = m.ax[2][iz + 1];
    const float cx[8] = {x0, x1, x1, x0, x0, x1, x1, x0};
    const float cy[8] = {y0, y0, y1, y1, y0, y0, y1, y1};
    const float cz[8] = {z0, z0, z0, z0, z1, z1, z1, z1};
    for (int q = 0; q < 8; ++q) {
        float a = 0, b = 0, c = 0;
        for (int k = 0; k < 8; ++k) {
            a += r.N[q][k] * cx[k];
            b += r.N[q][k] * cy[k];
            c += r.N[q][k] * cz[k];
        }
        xq[q] = a;
        yq[q] = b;
        zq[q] = c;
    }
}

// Coarse trilinear values and gradients at a point, in the point's own
// (clipped) coarse element (compute3DN cF:1361-1393; cF:1292-1342).
inline void coarse_shape(const Mesh& c, float x, float y, float z, int e[3], float N[8],
                         float dN[8][3]) {
    e[0] = coarse_elem(c, 0, x);
    e[1] = coarse_elem(c, 1, y);
    e[2] = coarse_elem(c, 2, z);
    const float x0 = c.ax[0][e[0]], x1 = c.ax[0][e[0] + 1];
    const float y0 = c.ax[1][e[1]], y1 = c.ax[1][e[1] + 1];
    const float z0 = c.ax[2][e[2]], z1 = c.ax[2][e[2] + 1];
    const float hv = c.h[0] * c.h[1] * c.h[2];
    const float inv = 1.0f / hv;
    const float px[2] = {x1 - x, x - x0}, py[2] = {y1 - y, y - y0}, pz[2] = {z1 - z, z - z0};
    static const int xb[8] = {0, 1, 1, 0, 0, 1, 1, 0}, yb[8] = {0, 0, 1, 1, 0, 0, 1, 1},
                     zb[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    for (int a = 0; a < 8; ++a) {
        N[a] = px[xb[a]] * py[yb[a]] * pz[zb[a]] * inv;
        if (dN) {
            dN[a][0] = (float)kXB[a] * py[yb[a]] * pz[zb[a]] / hv;
            dN[a][1] = px[xb[a]] * (float)kYB[a] * pz[zb[a]] / hv;
            dN[a][2] = px[xb[a]] * py[yb[a]] * (float)kZB[a] / hv;
        }
    }
}

}  // namespace

extern "C" {

int gmc_threads(void) { return omp_get_max_threads(); }

// solveMatrixFreeFE (cF:582-642):
//   out = [sum_e (M_e T_e - K_e T_e) + Fc + Corr] / sum_e M_e
// over the first `ne` elements; k_e, rhocp_e are the means of the 8 nodal
// values (cF:625-626).  Corr may be NULL (the scalar 0 of cF:3045).
int gmc_solve(const float* x, const float* y, const float* z, const int32_t* n, int64_t ne,
              const float* k, const float* rhocp, float dt, const float* T, const float* Fc,
              const float* Corr, float* out) {
    const Mesh m = mesh(x, y, z, n);
    const RefElem r = ref_element(m);
    float BTBw[8][8];
    for (int a = 0; a < 8; ++a)
        for (int b = 0; b < 8; ++b) {
            float s = 0;
            for (int q = 0; q < 8; ++q)
                for (int d = 0; d < 3; ++d) s += r.dN[q][a][d] * r.dN[q][b][d];
            BTBw[a][b] = s * r.w;
        }
    float Mrow[8];
    for (int a = 0; a < 8; ++a) {
        float s = 0;
        for (int b = 0; b < 8; ++b) {
            float ntn = 0;
            for (int q = 0; q < 8; ++q) ntn += r.N[q][b] * r.N[q][a];
            s += ntn * r.w;
        }
        Mrow[a] = s;
    }
    const int64_t nn = m.nn();
    // nodal sums, kept between calls (the port is driven from one thread)
    static std::vector<float> newT, newM;
    if ((int64_t)newT.size() < nn) {
        newT.resize(nn);
        newM.resize(nn);
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nn; ++i) newT[i] = newM[i] = 0.0f;
    for_rows_coloured(m, ne, [&](int iy, int iz, int nx_e) {
        int64_t id[8];
        for (int ix = 0; ix < nx_e; ++ix) {
            elem_nodes(m, ix, iy, iz, id);
            float ks = 0, cs = 0, Te[8];
            for (int a = 0; a < 8; ++a) {
                ks += k[id[a]];
                cs += rhocp[id[a]];
                Te[a] = T[id[a]];
            }
            const float kv = ks / 8.0f, mv = (cs / 8.0f) / dt;
            for (int a = 0; a < 8; ++a) {
                float kt = 0;
                for (int b = 0; b < 8; ++b) kt += Te[b] * BTBw[a][b];
                const float Me = mv * Mrow[a];
                newT[id[a]] += Me * Te[a] - kt * kv;
                newM[id[a]] += Me;
            }
        }
    });
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nn; ++i)
        out[i] = (newT[i] + Fc[i] + (Corr ? Corr[i] : 0.0f)) / newM[i];
    return 0;
}

// computeStateProperties (cF:2567-2614) with optional piecewise-linear
// tables for the bulk k and solid cp (linear extension past the end knots).
// mat: [T_sol, T_liq, k_powder, k_a0, k_a1, k_fluid, cp_a0, cp_a1, cp_mushy,
//       cp_fluid, rho]; kt/ct: knots then values (nk/nc each), nk/nc = 0
// for the linear law.  S1 (float 0/1) in, S1 out, S2 out (uint8).
int gmc_props(int64_t nn, const float* T, const float* S1in, int64_t n_sub, const double* mat,
              const double* kt, int nk, const double* ct, int nc, float* S1out, uint8_t* S2,
              float* k, float* rhocp) {
    const float Tsol = (float)mat[0], Tliq = (float)mat[1];
    auto table = [](double t, const double* tb, int n) {
        const double* kn = tb;
        const double* v = tb + n;
        int j = (int)(std::upper_bound(kn, kn + n, t) - kn) - 1;
        j = std::min(std::max(j, 0), n - 2);
        const double s = (v[j + 1] - v[j]) / (kn[j + 1] - kn[j]);
        return (float)(v[j] + s * (t - kn[j]));
    };
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nn; ++i) {
        const float t = T[i];
        const bool s2 = t >= Tliq, s3 = t > Tsol && t < Tliq;
        const float s1 = (i < n_sub || S1in[i] > 0.499f || s2) ? 1.0f : 0.0f;
        const float f2 = s2 ? 1.0f : 0.0f, f3 = s3 ? 1.0f : 0.0f;
        const float kb = nk ? table(t, kt, nk) : (float)mat[4] * t + (float)mat[3];
        k[i] = ((1 - s1) * (1 - f2) * (float)mat[2] + s1 * (1 - f2) * kb + f2 * (float)mat[5]) /
               1000.0f;
        const float cs = nc ? table(t, ct, nc) : (float)mat[7] * t + (float)mat[6];
        rhocp[i] = (float)mat[10] * ((1 - f2) * (1 - f3) * cs + f3 * (float)mat[8] +
                                     f2 * (float)mat[9]);
        S1out[i] = s1;
        S2[i] = s2;
    }
    return 0;
}

// computeSourcesL3 (cF:2960-3012): F_a = sum_e sum_q N_a(x_q) Q(x_q) w.
// src: [eta, radius, depth]; v: laser position (x, y, z).
int gmc_source_l3(const float* x, const float* y, const float* z, const int32_t* n,
                  const float* src, const float* v, float P, float* Fout) {
    const Mesh m = mesh(x, y, z, n);
    const RefElem r = ref_element(m);
    const Src s = src_consts(src);
    const int64_t nn = m.nn();
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nn; ++i) Fout[i] = 0.0f;
    const int64_t ne = (int64_t)m.ne[0] * m.ne[1] * m.ne[2];
    for_rows_coloured(m, ne, [&](int iy, int iz, int nx_e) {
        int64_t id[8];
        float xq[8], yq[8], zq[8], Qw[8];
        for (int ix = 0; ix < nx_e; ++ix) {
            gauss_xyz(m, r, ix, iy, iz, xq, yq, zq);
            for (int q = 0; q < 8; ++q) Qw[q] = source_q(s, P, v, xq[q], yq[q], zq[q]) * r.w;
            elem_nodes(m, ix, iy, iz, id);
            for (int a = 0; a < 8; ++a) {
                float f = 0;
                for (int q = 0; q < 8; ++q) f += r.N[a][q] * Qw[q];
                Fout[id[a]] += f;
            }
        }
    });
    return 0;
}

// computeLevelSource (cF:2667-2730) and the coarse projections of
// computeSources (cF:928-988): the fine-level (L3) source averaged over npos
// laser rows, projected on the coarse shape functions at the fine Gauss
// points (each point's values in its own clipped coarse element) and added to
// the coarse nodes of the coarse element holding Gauss point 0 (the BCOO of
// cF:1354-1357).  rows: npos x 7 (x, y, z, ..., dt, P); P per row.
int gmc_level_source(const float* cx, const float* cy, const float* cz, const int32_t* cn,
                     const float* fx, const float* fy, const float* fz, const int32_t* fn,
                     const float* src, const float* rows, const float* P, int npos,
                     float* Fout) {
    const Mesh c = mesh(cx, cy, cz, cn);
    const Mesh f = mesh(fx, fy, fz, fn);
    const RefElem r = ref_element(f);
    const Src s = src_consts(src);
    const int64_t nnc = c.nn();
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nnc; ++i) Fout[i] = 0.0f;
    const int64_t nef = (int64_t)f.ne[0] * f.ne[1] * f.ne[2];
    const float inv_np = 1.0f / (float)npos;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t e = 0; e < nef; ++e) {
        const int ix = (int)(e % f.ne[0]), iy = (int)((e / f.ne[0]) % f.ne[1]),
                  iz = (int)(e / ((int64_t)f.ne[0] * f.ne[1]));
        float xq[8], yq[8], zq[8];
        gauss_xyz(f, r, ix, iy, iz, xq, yq, zq);
        float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        int e0[3] = {0, 0, 0};
        for (int q = 0; q < 8; ++q) {
            float Nc[8];
            int ec[3];
            coarse_shape(c, xq[q], yq[q], zq[q], ec, Nc, nullptr);
            if (q == 0) memcpy(e0, ec, sizeof(e0));
            float Qs = 0;
            for (int p = 0; p < npos; ++p)
                Qs += source_q(s, P[p], rows + 7 * p, xq[q], yq[q], zq[q]) * r.w;
            for (int a = 0; a < 8; ++a) acc[a] += Nc[a] * Qs;
        }
        int64_t id[8];
        elem_nodes(c, e0[0], e0[1], e0[2], id);
        for (int a = 0; a < 8; ++a) atomic_add(&Fout[id[a]], acc[a] * inv_np);
    }
    return 0;
}

// Flux or mass restriction of fine T' terms onto a coarse level (the Shapes
// products of cF:1396-1565, 2733-2914, 3057-3221):
//   flux: V_A += -sum_e sum_q grad N_A . (k_e grad Tp) w      (mode 0)
//   mass: V_A += -sum_e sum_q N_A rhocp_e dTp w / dt          (mode 1)
// with k_e, rhocp_e the means over the Gauss points of the interpolated nodal
// values (cF:2773, 3106).  V is added to (full coarse node count).
int gmc_restrict(const float* cx, const float* cy, const float* cz, const int32_t* cn,
                 const float* fx, const float* fy, const float* fz, const int32_t* fn, int mode,
                 const float* u, const float* coef, float dt, float* V) {
    const Mesh c = mesh(cx, cy, cz, cn);
    const Mesh f = mesh(fx, fy, fz, fn);
    const RefElem r = ref_element(f);
    const int64_t nef = (int64_t)f.ne[0] * f.ne[1] * f.ne[2];
    const float wdt = mode ? (1.0f / dt) * r.w : r.w;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t e = 0; e < nef; ++e) {
        const int ix = (int)(e % f.ne[0]), iy = (int)((e / f.ne[0]) % f.ne[1]),
                  iz = (int)(e / ((int64_t)f.ne[0] * f.ne[1]));
        int64_t id[8];
        elem_nodes(f, ix, iy, iz, id);
        float ue[8], ce[8];
        for (int a = 0; a < 8; ++a) {
            ue[a] = u[id[a]];
            ce[a] = coef[id[a]];
        }
        float cm = 0;
        for (int q = 0; q < 8; ++q) {
            float s = 0;
            for (int a = 0; a < 8; ++a) s += r.N[q][a] * ce[a];
            cm += s;
        }
        cm /= 8.0f;
        float xq[8], yq[8], zq[8];
        gauss_xyz(f, r, ix, iy, iz, xq, yq, zq);
        float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        int e0[3] = {0, 0, 0};
        for (int q = 0; q < 8; ++q) {
            float Nc[8], dNc[8][3];
            int ec[3];
            coarse_shape(c, xq[q], yq[q], zq[q], ec, Nc, dNc);
            if (q == 0) memcpy(e0, ec, sizeof(e0));
            if (mode == 0) {
                float g[3] = {0, 0, 0};
                for (int a = 0; a < 8; ++a)
                    for (int d = 0; d < 3; ++d) g[d] += r.dN[q][a][d] * ue[a];
                for (int a = 0; a < 8; ++a)
                    for (int d = 0; d < 3; ++d) acc[a] += -dNc[a][d] * (cm * g[d]) * wdt;
            } else {
                float s = 0;
                for (int a = 0; a < 8; ++a) s += r.N[q][a] * ue[a];
                for (int a = 0; a < 8; ++a) acc[a] += -Nc[a] * (s * cm) * wdt;
            }
        }
        int64_t ic[8];
        elem_nodes(c, e0[0], e0[1], e0[2], ic);
        for (int a = 0; a < 8; ++a) atomic_add(&V[ic[a]], acc[a]);
    }
    return 0;
}

// computeConvRadBC (cF:2207-2301) on the top layer of the first `ne`
// elements.  mat: [T_amb, T_boil, h_conv, sigma*eps, evc, CP, CT, CM, L_ev,
// cp_fluid].  Fout = Fin + scatter (Fin may alias Fout).
int gmc_conv_rad(const float* x, const float* y, const float* z, const int32_t* n, int64_t ne,
                 const float* T, const float* mat, const float* Fin, float* Fout) {
    const Mesh m = mesh(x, y, z, n);
    const int64_t nn = m.nn();
    if (Fout != Fin) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nn; ++i) Fout[i] = Fin ? Fin[i] : 0.0f;
    }
    // 4-node face element at 2x2 Gauss (cF:778-855)
    const double g = 1.0 / sqrt(3.0);
    const double ki[4] = {-1, 1, 1, -1}, ei[4] = {-1, -1, 1, 1};
    float N[4][4];
    for (int q = 0; q < 4; ++q)
        for (int a = 0; a < 4; ++a)
            N[q][a] = (float)(0.25 * (1 + g * ki[q] * ki[a]) * (1 + g * ei[q] * ei[a]));
    const float w = (float)(0.25 * (double)m.h[0] * m.h[1]);
    const int64_t exy = (int64_t)m.ne[0] * m.ne[1];
    const int64_t iz = ne / exy - 1;  // top layer of the active elements
    if (iz < 0) return 0;
    const float Tamb = mat[0], Tb = mat[1], hc = mat[2], se = mat[3], evc = mat[4], CP = mat[5],
                CT = mat[6], CM = mat[7], Lev = mat[8], cpf = mat[9];
    const float Tamb4 = Tamb * Tamb * Tamb * Tamb;
    for (int c = 0; c < 2; ++c) {
        const int ny2 = (m.ne[1] - c + 1) / 2;
#pragma omp parallel for schedule(dynamic, 4)
        for (int r = 0; r < ny2; ++r) {
            const int iy = c + 2 * r;
            int64_t id[8];
            for (int ix = 0; ix < m.ne[0]; ++ix) {
                elem_nodes(m, ix, iy, (int)iz, id);
                const float Tt[4] = {T[id[4]], T[id[5]], T[id[6]], T[id[7]]};
                float qw[4];
                for (int q = 0; q < 4; ++q) {
                    float Tq = 0;
                    for (int a = 0; a < 4; ++a) Tq += N[q][a] * Tt[a];
                    Tq = std::min(Tq, Tb + 1000.0f);
                    const float invT = 1.0f / Tq;
                    const float Epv = Lev + cpf * (Tq - Tamb);
                    const float mm = sqrtf(CM * invT);
                    const float S = evc * CP * expf(-CT * (invT - 1.0f / Tb)) * mm * Epv;
                    float fl = hc * (Tamb - Tq) + se * (Tamb4 - Tq * Tq * Tq * Tq) - S;
                    qw[q] = fl * 1e-6f * w;
                }
                for (int a = 0; a < 4; ++a) {
                    float s = 0;
                    for (int q = 0; q < 4; ++q) s += N[q][a] * qw[q];
                    Fout[id[4 + a]] += s;
                }
            }
        }
    }
    return 0;
}

// interpolatePoints (cF:1131-1210): trilinear interpolation of a nodal field
// onto the grid of target axes tx, ty, tz (x fastest); element index clipped,
// all weights zeroed when one leaves [-0.01, 1.01], else clipped to [0, 1].
int gmc_interp(const float* x, const float* y, const float* z, const int32_t* n, const float* u,
               const float* tx, const float* ty, const float* tz, const int32_t* tn,
               float* out) {
    const Mesh m = mesh(x, y, z, n);
    const int64_t nt = (int64_t)tn[0] * tn[1] * tn[2];
    const float inv = 1.0f / (m.h[0] * m.h[1] * m.h[2]);
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < nt; ++t) {
        const int i = (int)(t % tn[0]), j = (int)((t / tn[0]) % tn[1]),
                  k = (int)(t / ((int64_t)tn[0] * tn[1]));
        const float p[3] = {tx[i], ty[j], tz[k]};
        int e[3];
        for (int d = 0; d < 3; ++d) e[d] = coarse_elem(m, d, p[d]);
        const float a[2] = {m.ax[0][e[0] + 1] - p[0], p[0] - m.ax[0][e[0]]};
        const float b[2] = {m.ax[1][e[1] + 1] - p[1], p[1] - m.ax[1][e[1]]};
        const float c[2] = {m.ax[2][e[2] + 1] - p[2], p[2] - m.ax[2][e[2]]};
        static const int xb[8] = {0, 1, 1, 0, 0, 1, 1, 0}, yb[8] = {0, 0, 1, 1, 0, 0, 1, 1},
                         zb[8] = {0, 0, 0, 0, 1, 1, 1, 1};
        float Nw[8];
        bool ok = true;
        for (int q = 0; q < 8; ++q) {
            Nw[q] = a[xb[q]] * b[yb[q]] * c[zb[q]] * inv;
            ok = ok && Nw[q] >= -1e-2f && Nw[q] <= 1.0f + 1e-2f;
        }
        int64_t id[8];
        elem_nodes(m, e[0], e[1], e[2], id);
        float s = 0;
        if (ok)
            for (int q = 0; q < 8; ++q) s += std::min(std::max(Nw[q], 0.0f), 1.0f) * u[id[q]];
        out[t] = s;
    }
    return 0;
}

}  // extern "C"
