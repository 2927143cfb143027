// Host side of libgomelt.so: the C ABI of include/gomelt.h and the engine
// that drives one L1 step of the multilevel scheme on a HIP stream.
//
// The engine owns the device state of the three levels and of the Level0
// state grid.  The host keeps the window bookkeeping (integer shifts, overlap
// boxes) and writes it, with the toolpath rows of the step, into a GmStepDesc
// that is uploaded once per call; every kernel reads the moving meshes from
// that descriptor, so the ~330 launches of a subcycled L1 step are recorded
// once into a hipGraph and replayed for every later step.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "gm_internal.h"

using namespace gm;

static thread_local std::string g_err;

static int fail(const std::string& msg) {
    g_err = msg;
    return -1;
}

#define HIPCHK(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            (void)hipGetLastError(); /* the caller gets the error; do not leave it */   \
            return fail(std::string(#call) + ": " + hipGetErrorString(e_));             \
        }                                                                               \
    } while (0)

extern "C" const char* gm_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// Device, memory, streams
// ---------------------------------------------------------------------------

static int scratch_ptr(void** out);

extern "C" int gm_device_init(int device) {
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipFree(nullptr));
    void* p = nullptr;
    return scratch_ptr(&p);  // the device's address of the no-allocation scratch
}

extern "C" int gm_device_name(char* buf, int len) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, dev));
    snprintf(buf, len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return 0;
}

extern "C" int gm_malloc(void** ptr, size_t bytes) {
    HIPCHK(hipMalloc(ptr, bytes ? bytes : 1));
    return 0;
}
extern "C" int gm_free(void* ptr) {
    HIPCHK(hipFree(ptr));
    return 0;
}
extern "C" int gm_memcpy_h2d(void* dst, const void* src, size_t n, void* s) {
    HIPCHK(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, (hipStream_t)s));
    return 0;
}
extern "C" int gm_memcpy_d2h(void* dst, const void* src, size_t n, void* s) {
    HIPCHK(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, (hipStream_t)s));
    HIPCHK(hipStreamSynchronize((hipStream_t)s));
    return 0;
}
extern "C" int gm_memcpy_d2d(void* dst, const void* src, size_t n, void* s) {
    HIPCHK(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToDevice, (hipStream_t)s));
    return 0;
}
extern "C" int gm_memset(void* dst, int v, size_t n, void* s) {
    HIPCHK(hipMemsetAsync(dst, v, n, (hipStream_t)s));
    return 0;
}
extern "C" int gm_stream_create(void** s) {
    HIPCHK(hipStreamCreate((hipStream_t*)s));
    return 0;
}
extern "C" int gm_stream_destroy(void* s) {
    HIPCHK(hipStreamDestroy((hipStream_t)s));
    return 0;
}
extern "C" int gm_stream_sync(void* s) {
    HIPCHK(hipStreamSynchronize((hipStream_t)s));
    HIPCHK(hipGetLastError());
    return 0;
}
extern "C" int gm_event_create(void** ev) {
    HIPCHK(hipEventCreate((hipEvent_t*)ev));
    return 0;
}
extern "C" int gm_event_destroy(void* ev) {
    HIPCHK(hipEventDestroy((hipEvent_t)ev));
    return 0;
}
extern "C" int gm_event_record(void* ev, void* s) {
    HIPCHK(hipEventRecord((hipEvent_t)ev, (hipStream_t)s));
    return 0;
}
extern "C" int gm_stream_wait_event(void* s, void* ev) {
    HIPCHK(hipStreamWaitEvent((hipStream_t)s, (hipEvent_t)ev, 0));
    return 0;
}
extern "C" int gm_event_elapsed_ms(void* a, void* b, float* ms) {
    HIPCHK(hipEventSynchronize((hipEvent_t)b));
    HIPCHK(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
    return 0;
}

// ---------------------------------------------------------------------------
// Helpers shared by the single operators and the engine
// ---------------------------------------------------------------------------

static int64_t grid_nn(const gm_grid& g) { return (int64_t)g.n[0] * g.n[1] * g.n[2]; }

// Truncation toward zero of (v/h + 0.01) in FP32 (cF:1716-1718).
static int shift_of(float v, float h) { return (int)truncf(v / h + 0.01f); }

// Bound on the coarse elements a fine mesh can cover along one axis.
static int cover_bound(const gm_grid& c, const gm_grid& f, int d) {
    const int nec = c.n[d] - 1;
    const double ext = (double)(f.n[d] - 1) * f.h[d];
    const int b = (int)ceil(ext / c.h[d]) + 2;
    return std::max(1, std::min(nec, b));
}

struct PairMem {
    GmPair p{};
    void* blob = nullptr;
};

// Sizes of a fine -> coarse pair and the bytes of its axis tables.
static size_t pair_layout(GmPair& p, int coarse, int fine, const gm_grid& gc, const gm_grid& gf) {
    p.coarse = coarse;
    p.fine = fine;
    size_t bytes = 0;
    for (int d = 0; d < 3; ++d) {
        p.nef[d] = gf.n[d] - 1;
        p.cmax[d] = cover_bound(gc, gf, d);
        bytes += (size_t)p.nef[d] * 4 + (size_t)p.nef[d] * 2 * 16 + (size_t)(2 + 2 * p.cmax[d]) * 4;
        bytes += 64;
    }
    return bytes;
}

// Point the axis tables of a laid-out pair into `blob` (64-byte aligned).
static void pair_bind(GmPair& p, void* blob) {
    char* b = (char*)blob;
    for (int d = 0; d < 3; ++d) {
        p.ax[d].q = (float4*)b;
        b += (size_t)p.nef[d] * 2 * 16;
        p.ax[d].c0 = (int32_t*)b;
        b += (size_t)p.nef[d] * 4;
        p.ax[d].info = (int32_t*)b;
        b += (size_t)(2 + 2 * p.cmax[d]) * 4;
        b = (char*)(((uintptr_t)b + 63) & ~(uintptr_t)63);
    }
}

static int pair_alloc(PairMem& pm, int coarse, int fine, const gm_grid& gc, const gm_grid& gf) {
    const size_t bytes = pair_layout(pm.p, coarse, fine, gc, gf);
    HIPCHK(hipMalloc(&pm.blob, bytes));
    HIPCHK(hipMemset(pm.blob, 0, bytes));
    pair_bind(pm.p, pm.blob);
    return 0;
}

static int64_t pair_nodes(const GmPair& p) {
    return (int64_t)(p.cmax[0] + 1) * (p.cmax[1] + 1) * (p.cmax[2] + 1);
}
static int64_t pair_elems(const GmPair& p) { return (int64_t)p.cmax[0] * p.cmax[1] * p.cmax[2]; }
static int pair_gmax(const GmPair& p) {
    return std::max(p.cmax[0], std::max(p.cmax[1], p.cmax[2])) + 1;
}

static LSArgs ls_default() {
    LSArgs a;
    memset(&a, 0, sizeof(a));
    a.use_desc = 1;
    a.dt_mode = GM_DT_VALUE;
    a.src_row = -1;
    a.alpha = 1.0f;
    return a;
}

static void ls_coeffs(LSArgs& a, const gm_grid& g) { level_coeffs(g, a.kc, &a.mrow, &a.w2, &a.w3); }

// ---------------------------------------------------------------------------
// Single operators
// ---------------------------------------------------------------------------

extern "C" int gm_state_properties(const gm_material* mat, const float* T, const float* S1,
                                   int64_t nn, int64_t nsub, float* S1o, uint8_t* S2o, float* k,
                                   float* c, void* stream) {
    launch_props(make_devmat(*mat), T, S1, nn, nsub, S1o, S2o, k, c, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_solve_matrix_free(const gm_grid* grid, int64_t ne_active, const float* k,
                                    const float* rhocp, float dt, const float* T, const float* F,
                                    const float* Corr, float* T_out, void* stream) {
    LSArgs a = ls_default();
    a.use_desc = 0;
    a.grid = *grid;
    a.T = T;
    a.kin = k;
    a.cin = rhocp;
    a.Tout = T_out;
    a.ne_active = ne_active;
    a.nn_active = grid_nn(*grid);
    a.dt = dt;
    a.Fext = F;
    a.corr_full = Corr;
    ls_coeffs(a, *grid);
    // the given-k path reads only T_amb of the material (zero here: every
    // node is active)
    gm_material dummy;
    memset(&dummy, 0, sizeof(dummy));
    const DevMat hm = make_devmat(dummy);
    DevMat* dm = nullptr;
    HIPCHK(hipMalloc((void**)&dm, sizeof(DevMat)));
    hipError_t rc = hipMemcpyAsync(dm, &hm, sizeof(DevMat), hipMemcpyHostToDevice,
                                   (hipStream_t)stream);
    if (rc == hipSuccess) {
        launch_level_step(a, nullptr, *grid, dm, (hipStream_t)stream);
        rc = hipGetLastError();
    }
    if (rc == hipSuccess) rc = hipStreamSynchronize((hipStream_t)stream);
    (void)hipFree(dm);
    HIPCHK(rc);
    return 0;
}

extern "C" int gm_source_l3(const gm_grid* g, const gm_material* mat, const float* v, float P,
                            float* F, void* stream) {
    launch_source_l3_full(*g, make_devmat(*mat), v[0], v[1], v[2], P, F, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_source_l3_mean(const gm_grid* g, const gm_material* mat,
                                 const float* rows_host, int n_rows, float* F, void* stream) {
    if (!g || !mat || !rows_host || !F || n_rows < 1)
        return fail("gm_source_l3_mean: null argument or no rows");
    launch_source_l3_mean(*g, make_devmat(*mat), rows_host, n_rows, F, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_conv_rad_bc(const gm_grid* g, const gm_material* mat, const float* T,
                              int64_t ne_active, float* F, void* stream) {
    launch_convrad_full(*g, make_devmat(*mat), T, ne_active, F, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_melt_pool(const gm_grid* g, const float* T, float T_liquidus, float* dims,
                            void* stream) {
    if (!g || !dims || !T) return fail("gm_melt_pool: null argument");
    if (g->n[0] < 1 || g->n[1] < 1 || g->n[2] < 1 ||
        (int64_t)g->n[0] * g->n[1] * g->n[2] >= (int64_t)1 << 31)
        return fail("gm_melt_pool: node counts must be >= 1 and below 2^31 in total");
    launch_melt_pool(*g, T, T_liquidus, dims, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_energy_balance(const gm_grid* g, int64_t ne_active, const float* k,
                                 const float* rhocp, float dt, const float* T, const float* F,
                                 const float* Corr, const float* T_new, int bc_faces,
                                 float T_clamp, double* sums, void* stream) {
    if (!g || !k || !rhocp || !T || !T_new || !sums)
        return fail("gm_energy_balance: null argument");
    if (g->n[0] < 2 || g->n[1] < 2 || g->n[2] < 2 ||
        (int64_t)g->n[0] * g->n[1] * g->n[2] >= (int64_t)1 << 31)
        return fail("gm_energy_balance: need >= 2 nodes per axis and < 2^31 nodes");
    if (bc_faces < 0 || bc_faces > 31) return fail("gm_energy_balance: bc_faces is a 5-bit mask");
    launch_energy(*g, ne_active, k, rhocp, T, F, Corr, T_new, bc_faces, T_clamp, dt, sums,
                  (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_interpolate_points(const gm_grid* src, const float* u, const gm_grid* tgt,
                                     float* out, void* stream) {
    launch_interp_grid(nullptr, -1, src, u, nullptr, -1, tgt, grid_nn(*tgt), out,
                       (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_interpolate_matrix(const gm_grid* src, const gm_grid* tgt, float* Nc,
                                     int32_t* node, void* stream) {
    launch_interp_matrix(*src, *tgt, Nc, node, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_interpolate_w_matrix(const float* Nc, const int32_t* node, int64_t n,
                                       const float* T, float* out, void* stream) {
    launch_interp_w(Nc, node, n, T, out, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

// A temporary descriptor holding explicit grids, for the single operators
// whose kernels read meshes from a descriptor.
struct TmpDesc {
    GmStepDesc* d = nullptr;
    ~TmpDesc() {
        if (d) (void)hipFree(d);
    }
};

static int tmp_desc(TmpDesc& t, const gm_grid* g1, const gm_grid* g2, const gm_grid* g3,
                    const GmStepDesc* extra, hipStream_t s) {
    GmStepDesc h;
    if (extra) h = *extra;
    else memset(&h, 0, sizeof(h));
    if (g1) h.g[1] = *g1;
    if (g2) h.g[2] = *g2;
    if (g3) h.g[3] = *g3;
    HIPCHK(hipMalloc(&t.d, sizeof(GmStepDesc)));
    HIPCHK(hipMemcpyAsync(t.d, &h, sizeof(GmStepDesc), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

int gm_eval_tables(const LSArgs& a, const gm_grid& g, float* F, hipStream_t s);

extern "C" int gm_level_source(const gm_grid* coarse, const gm_grid* fine, const gm_material* mat,
                               const float* rows, int npos, float* F_out, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    TmpDesc td;
    if (tmp_desc(td, coarse, fine, nullptr, nullptr, s)) return -1;
    PairMem pm;
    if (pair_alloc(pm, 1, 2, *coarse, *fine)) return -1;
    float* rows_d = nullptr;
    float* G = nullptr;
    const int nmax = pair_gmax(pm.p);
    HIPCHK(hipMalloc(&rows_d, (size_t)npos * 7 * 4));
    HIPCHK(hipMalloc(&G, (size_t)npos * 3 * nmax * 4));
    HIPCHK(hipMemcpyAsync(rows_d, rows, (size_t)npos * 7 * 4, hipMemcpyHostToDevice, s));
    launch_axis_tables(&pm.p, 1, td.d, s);
    launch_source_tables(pm.p, td.d, make_devmat(*mat), 0, npos, 1, rows_d, G, nmax, s);
    LSArgs a = ls_default();
    a.G = G;
    a.G_npos = npos;
    a.G_nmax = nmax;
    float kc[8], mrow, w2, w3f;
    level_coeffs(*fine, kc, &mrow, &w2, &w3f);
    a.G_scale = w3f / (float)npos;
    for (int d = 0; d < 3; ++d) a.G_info[d] = pm.p.ax[d].info;
    const int rc = gm_eval_tables(a, *coarse, F_out, s);
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipFree(rows_d));
    HIPCHK(hipFree(G));
    HIPCHK(hipFree(pm.blob));
    return rc;
}

// Evaluate F_A = G_scale * sum_pos Gx Gy Gz on a full coarse array.
__global__ void k_eval_tables(LSArgs a, gm_grid g, float* F) {
    const int64_t n = (int64_t)g.n[0] * g.n[1] * g.n[2];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int ix = t % g.n[0], iy = (t / g.n[0]) % g.n[1], iz = t / ((int64_t)g.n[0] * g.n[1]);
    const int bx = ix - a.G_info[0][0], by = iy - a.G_info[1][0], bz = iz - a.G_info[2][0];
    float v = 0.0f;
    if (bx >= 0 && bx <= a.G_info[0][1] && by >= 0 && by <= a.G_info[1][1] && bz >= 0 &&
        bz <= a.G_info[2][1]) {
        float s = 0.0f;
        for (int p = 0; p < a.G_npos; ++p) {
            const float* Gp = a.G + (int64_t)p * 3 * a.G_nmax;
            s += Gp[bx] * Gp[a.G_nmax + by] * Gp[2 * a.G_nmax + bz];
        }
        v = s * a.G_scale;
    }
    F[t] = v;
}

int gm_eval_tables(const LSArgs& a, const gm_grid& g, float* F, hipStream_t s) {
    const int64_t n = grid_nn(g);
    hipLaunchKernelGGL(k_eval_tables, dim3((n + 255) / 256), dim3(256), 0, s, a, g, F);
    HIPCHK(hipGetLastError());
    return 0;
}

// Expand a node-box array of a pair onto the full coarse array.
__global__ void k_box_to_full(GmPair p, const float* __restrict__ V, gm_grid g, float* out) {
    const int64_t n = (int64_t)g.n[0] * g.n[1] * g.n[2];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int ix = t % g.n[0], iy = (t / g.n[0]) % g.n[1], iz = t / ((int64_t)g.n[0] * g.n[1]);
    const int bx = ix - p.ax[0].info[0], by = iy - p.ax[1].info[0], bz = iz - p.ax[2].info[0];
    const int mx = p.cmax[0] + 1, my = p.cmax[1] + 1;
    float v = 0.0f;
    if (bx >= 0 && bx <= p.ax[0].info[1] && by >= 0 && by <= p.ax[1].info[1] && bz >= 0 &&
        bz <= p.ax[2].info[1])
        v = V[bx + mx * (by + my * bz)];
    out[t] += v;
}

extern "C" int gm_restrict(const gm_grid* coarse, const gm_grid* fine, const gm_material* mat,
                           const float* T, const float* S1, int64_t nsub, const float* Tp,
                           const float* TpA, const float* TpB, float dt, float* V_out,
                           void* stream) {
    hipStream_t s = (hipStream_t)stream;
    TmpDesc td;
    if (tmp_desc(td, coarse, fine, nullptr, nullptr, s)) return -1;
    PairMem pm;
    if (pair_alloc(pm, 1, 2, *coarse, *fine)) return -1;
    double* part = nullptr;
    float* V = nullptr;
    HIPCHK(hipMalloc(&part, pair_elems(pm.p) * 8 * sizeof(double)));
    HIPCHK(hipMalloc(&V, pair_nodes(pm.p) * 4));
    launch_axis_tables(&pm.p, 1, td.d, s);
    RArgs r;
    memset(&r, 0, sizeof(r));
    r.T = T;
    r.S1f = S1;
    r.n_sub = nsub;
    r.Tp = Tp;
    r.TpA = TpA;
    r.TpB = TpB;
    r.dt = dt;
    r.dt_mode = GM_DT_VALUE;
    float kc[8], mrow, w2, w3;
    level_coeffs(*fine, kc, &mrow, &w2, &w3);
    r.wf = w3;
    launch_restrict(pm.p, td.d, make_devmat(*mat), r, part, s);
    launch_restrict_nodes(pm.p, part, nullptr, nullptr, nullptr, V, nullptr, nullptr, 0, 0, 0.0f,
                          s);
    const int64_t n = grid_nn(*coarse);
    hipLaunchKernelGGL(k_box_to_full, dim3((n + 255) / 256), dim3(256), 0, s, pm.p, V, *coarse,
                       V_out);
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipFree(part));
    HIPCHK(hipFree(V));
    HIPCHK(hipFree(pm.blob));
    return 0;
}

// ---- gm_restrict without allocation ---------------------------------------
// Scratch of gm_restrict_noalloc: a step descriptor holding the two grids, the
// pair's axis tables, the FP64 element partials and the coarse node box, laid
// out one after the other (256-byte aligned) in one static device array.
// Every GPU holds its own copy until the process ends; host code reaches it
// only through the per-device address from hipGetSymbolAddress.
#define GM_NOALLOC_SCRATCH_BYTES ((size_t)16 << 20)
#define GM_NOALLOC_MAX_DEVICES 64
static __device__ __align__(256) uint8_t scratch[GM_NOALLOC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void* g_scratch_dev[GM_NOALLOC_MAX_DEVICES];

static int scratch_ptr(void** out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= GM_NOALLOC_MAX_DEVICES) return fail("device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return 0;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

__global__ void k_put_grids12(GmStepDesc* d, gm_grid g1, gm_grid g2) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        d->g[1] = g1;
        d->g[2] = g2;
    }
}

__global__ void k_zero_words(uint32_t* p, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0u;
}

extern "C" int gm_restrict_noalloc(const gm_grid* coarse, const gm_grid* fine,
                                   const gm_material* mat, const float* T, const float* S1,
                                   int64_t nsub, const float* Tp, const float* TpA,
                                   const float* TpB, float dt, float* V_out, void* stream) {
    if (!coarse || !fine || !mat || !T || !S1 || !V_out)
        return fail("gm_restrict_noalloc: null argument");
    for (int d = 0; d < 3; ++d)
        if (coarse->n[d] < 2 || fine->n[d] < 2)
            return fail("gm_restrict_noalloc: need >= 2 nodes per axis on both grids");
    hipStream_t s = (hipStream_t)stream;
    void* base = nullptr;
    if (scratch_ptr(&base)) return -1;
    GmPair p{};
    const size_t blob_bytes = pair_layout(p, 1, 2, *coarse, *fine);
    const size_t o_blob = align256(sizeof(GmStepDesc));
    const size_t o_part = o_blob + align256(blob_bytes);
    const size_t o_V = o_part + align256((size_t)pair_elems(p) * 8 * sizeof(double));
    const size_t total = o_V + align256((size_t)pair_nodes(p) * 4);
    if (total > GM_NOALLOC_SCRATCH_BYTES)
        return fail("gm_restrict_noalloc: this pair needs " + std::to_string(total) +
                    " bytes of scratch; the limit is " +
                    std::to_string(GM_NOALLOC_SCRATCH_BYTES) + " (use gm_restrict)");
    char* b = (char*)base;
    GmStepDesc* d = (GmStepDesc*)b;
    pair_bind(p, b + o_blob);
    double* part = (double*)(b + o_part);
    float* V = (float*)(b + o_V);
    // the descriptor's grids and cleared axis tables, written by kernels (no
    // host copy, no memset): the scratch keeps the previous call's contents
    hipLaunchKernelGGL(k_put_grids12, dim3(1), dim3(64), 0, s, d, *coarse, *fine);
    const int64_t words = (int64_t)(blob_bytes / 4);
    hipLaunchKernelGGL(k_zero_words, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, s,
                       (uint32_t*)(b + o_blob), words);
    launch_axis_tables(&p, 1, d, s);
    RArgs r;
    memset(&r, 0, sizeof(r));
    r.T = T;
    r.S1f = S1;
    r.n_sub = nsub;
    r.Tp = Tp;
    r.TpA = TpA;
    r.TpB = TpB;
    r.dt = dt;
    r.dt_mode = GM_DT_VALUE;
    float kc[8], mrow, w2, w3;
    level_coeffs(*fine, kc, &mrow, &w2, &w3);
    r.wf = w3;
    launch_restrict(p, d, make_devmat(*mat), r, part, s);
    launch_restrict_nodes(p, part, nullptr, nullptr, nullptr, V, nullptr, nullptr, 0, 0, 0.0f, s);
    const int64_t n = grid_nn(*coarse);
    hipLaunchKernelGGL(k_box_to_full, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, V,
                       *coarse, V_out);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_new_tprime(const gm_grid* fine, const float* FineT, const gm_grid* coarse,
                             float* CoarseT, const int32_t* ov_lo, const int32_t* ov_n, float* Tp,
                             void* stream) {
    hipStream_t s = (hipStream_t)stream;
    GmStepDesc h;
    memset(&h, 0, sizeof(h));
    for (int d = 0; d < 3; ++d) {
        h.ov3.lo[d] = ov_lo[d];
        h.ov3.n[d] = ov_n[d];
        h.ov3_xyz.n[d] = ov_n[d];
        h.ov3_xyz.x0[d] = coarse->x0[d] + coarse->h[d] * (float)ov_lo[d];
        h.ov3_xyz.h[d] = coarse->h[d];
    }
    TmpDesc td;
    if (tmp_desc(td, nullptr, coarse, fine, &h, s)) return -1;
    launch_inject(td.d, 3, 2, FineT, CoarseT, (int64_t)ov_n[0] * ov_n[1] * ov_n[2], s);
    launch_tprime(td.d, 3, 2, FineT, CoarseT, Tp, grid_nn(*fine), s);
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

// ---------------------------------------------------------------------------
// The engine
// ---------------------------------------------------------------------------

// Loop state of a subcycled L1 step run stage by stage (see stage_prologue).
struct SubcycleState {
    const float* T2cur;
    const uint8_t* S12cur;
    const float* T3cur;
    const float* Tp3cur;
    const uint8_t* S13cur;
    const uint8_t* S23cur;
    int b2, b3;
    // outputs of the current substep's L2 stage
    float* T2new;
    uint8_t* S12new;
    float* T3out;  // where the N3 plain solves leave the substep's final L3 field
    float* T3ext;  // where a caller writes its own final L3 field (never T3cur)
    const float* T3last;  // the final L3 field of the last closed substep
};

struct gm_engine {
    gm_config cfg;
    gm_material mat;
    DevMat dm;
    DevMat* d_dm = nullptr;  // device copy read by the level steps
    hipStream_t s;
    int64_t nn[4];
    int N2, N3, N23;
    // level operator constants
    float kc[4][8], mrow[4], w2[4], w3[4];
    // state
    float *T1 = nullptr, *S11 = nullptr, *S1st = nullptr;
    int nst = 0;
    float *T2 = nullptr, *Tp2 = nullptr;
    uint8_t* S12 = nullptr;
    float *T3 = nullptr, *Tp3 = nullptr;
    uint8_t *S13 = nullptr, *S23 = nullptr;
    uint8_t* L0S1 = nullptr;
    float *acc = nullptr, *macc = nullptr;
    uint8_t* s2snap = nullptr;
    int32_t* s2box = nullptr;
    // scratch
    float *T1p = nullptr, *T1c = nullptr;
    float *V1 = nullptr, *V1m = nullptr, *V2 = nullptr;
    float *G1 = nullptr, *G2 = nullptr;
    // coarse source on the node boxes: per L2 substep (pair 2), and a scratch
    // box for the node finishing passes
    float *F2 = nullptr, *Fg = nullptr;
    float* T2b[2] = {nullptr, nullptr};
    uint8_t* S12b[2] = {nullptr, nullptr};
    float* Tp2x = nullptr;
    std::vector<float*> L2all, L3all, histTp;
    float* T3b[2] = {nullptr, nullptr};
    float* T3x = nullptr;  // a caller's L3 jump of a Part-1 substep whose plain final
                           // buffer is its start buffer (even N3)
    uint8_t* S13b[2] = {nullptr, nullptr};
    uint8_t* S23b[2] = {nullptr, nullptr};
    float* Tp3x = nullptr;
    float* src3 = nullptr;  // L3 source factors of the N2*N3 rows of a subcycle
    float *acc3 = nullptr, *macc3 = nullptr;  // melt time on the L3 nodes during Part 2
    // k, cp companions of the L2/L3 temperature buffers that feed level
    // steps: kc_T[i] belongs to kc_key[i]
    std::vector<const float*> kc_key;
    std::vector<float*> kc_T;
    bool use_kc = false;      // restrictions read the carried k, cp
    bool fast_restrict = false;  // all pairs node-aligned: node-centric restriction
    int strip_R0[3] = {-1, -1, -1}, strip_a0[3] = {-1, -1, -1};  // restriction strip layout
    float mcs[4][8];          // consistent mass per neighbour pattern (x rho)
    uint8_t* preS2 = nullptr;
    uint8_t* reset3 = nullptr;  // L3 nodes whose S2 went 0 -> 1 in the last single step
    bool acc_external = false;  // melt time kept by the caller (window arrays acc3/macc3)
    float *fa2 = nullptr, *fb2 = nullptr, *fa3 = nullptr, *fb3 = nullptr;
    float* mm = nullptr;
    PairMem pair[3];
    double* part[3] = {nullptr, nullptr, nullptr};
    int gmax[3];
    // descriptor
    GmStepDesc host;
    GmStepDesc* d_desc = nullptr;
    GmStepDesc* ring[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t ring_ev[4];
    int ring_k = 0;
    // window state
    int s3[3] = {0, 0, 0}, s2[3] = {0, 0, 0};
    // staged subcycle (gm_engine_stage_*): state, and the next stage expected
    SubcycleState stg{};
    int stg_active = 0, stg_part = 0, stg_k = 0, stg_phase = 0;
    // graph
    bool use_graph = true;
    bool graph_ok = false;
    hipGraphExec_t gexec = nullptr;
    std::vector<void*> allocs;
};

static int eng_alloc(gm_engine* e, void** p, size_t bytes) {
    HIPCHK(hipMalloc(p, bytes ? bytes : 1));
    e->allocs.push_back(*p);
    return 0;
}

#define EALLOC(ptr, bytes)                                                     \
    do {                                                                       \
        if (eng_alloc(e, (void**)&(ptr), (size_t)(bytes))) return -1;          \
    } while (0)

static int upload_desc(gm_engine* e, size_t bytes) {
    e->ring_k = (e->ring_k + 1) & 3;
    HIPCHK(hipEventSynchronize(e->ring_ev[e->ring_k]));
    memcpy(e->ring[e->ring_k], &e->host, bytes);
    HIPCHK(hipMemcpyAsync(e->d_desc, e->ring[e->ring_k], bytes, hipMemcpyHostToDevice, e->s));
    HIPCHK(hipEventRecord(e->ring_ev[e->ring_k], e->s));
    return 0;
}

static size_t desc_bytes(int nrows) {
    return offsetof(GmStepDesc, rows) + (size_t)std::max(nrows, 1) * 7 * sizeof(float);
}

extern "C" int gm_engine_destroy(gm_engine* e) {
    if (!e) return 0;
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    for (void* p : e->allocs) (void)hipFree(p);
    for (int i = 0; i < 3; ++i)
        if (e->pair[i].blob) (void)hipFree(e->pair[i].blob);
    for (int i = 0; i < 4; ++i) {
        if (e->ring[i]) (void)hipHostFree(e->ring[i]);
        (void)hipEventDestroy(e->ring_ev[i]);
    }
    delete e;
    return 0;
}

static bool pairs_aligned(gm_engine* e);
static int update_alignment(gm_engine* e);
static void strip_phase(gm_engine* e, int i, int& R0, int& a0);

extern "C" int gm_engine_create(const gm_config* cfg, const gm_material* mat, void* stream,
                                gm_engine** out) {
    gm_engine* e = new gm_engine();
    *out = nullptr;
    e->cfg = *cfg;
    e->mat = *mat;
    e->dm = make_devmat(*mat);
    e->s = (hipStream_t)stream;
    e->N2 = cfg->n2;
    e->N3 = cfg->n3;
    e->N23 = cfg->n2 * cfg->n3;
    if (e->N23 > GM_MAX_ROWS || e->N2 > GM_MAX_N2) {
        delete e;
        return fail("subcycle counts exceed the engine limits (N2 <= 64, N2*N3 <= 1024)");
    }
    for (int l = 0; l < 4; ++l) e->nn[l] = grid_nn(cfg->level[l]);
    // the L2/L3 subcycle step addresses k|cp companions with 32-bit offsets
    if (e->nn[2] >= ((int64_t)1 << 28) || e->nn[3] >= ((int64_t)1 << 28)) {
        delete e;
        return fail("Level2/Level3 meshes are limited to 2^28 nodes");
    }
    // the level-step kernel takes k in W/(m K) and cp in J/(kg K): the 1/1000
    // of cF:2602 and rho of cF:2612 go into the stencil and mass constants
    for (int l = 1; l < 4; ++l) {
        level_coeffs(cfg->level[l], e->kc[l], &e->mrow[l], &e->w2[l], &e->w3[l]);
        for (int p = 0; p < 8; ++p) e->kc[l][p] = (float)((double)e->kc[l][p] / 1000.0);
        e->mrow[l] = (float)((double)e->mrow[l] * mat->rho);
        // consistent mass of a box element, per neighbour pattern: rho *
        // prod_d h_d (2 or 1)/6 (the 1-D mass matrix (h/6)[2 1; 1 2])
        for (int p = 0; p < 8; ++p) {
            double v = mat->rho;
            for (int d = 0; d < 3; ++d)
                v *= (double)cfg->level[l].h[d] * (((p >> d) & 1) ? 1.0 : 2.0) / 6.0;
            e->mcs[l][p] = (float)v;
        }
    }
    const int64_t n1 = e->nn[1], n2 = e->nn[2], n3 = e->nn[3], n0 = e->nn[0];
    e->nst = std::max(1, (int)lround((double)cfg->level[1].h[2] / cfg->layer_height));
    auto bad = [&](int rc) {
        gm_engine_destroy(e);
        return rc;
    };
    if (eng_alloc(e, (void**)&e->T1, n1 * 4) || eng_alloc(e, (void**)&e->S11, n1 * 4) ||
        eng_alloc(e, (void**)&e->S1st, (size_t)e->nst * n1 * 4) ||
        eng_alloc(e, (void**)&e->T2, n2 * 4) || eng_alloc(e, (void**)&e->Tp2, n2 * 4) ||
        eng_alloc(e, (void**)&e->S12, n2) || eng_alloc(e, (void**)&e->T3, n3 * 4) ||
        eng_alloc(e, (void**)&e->Tp3, n3 * 4) || eng_alloc(e, (void**)&e->S13, n3) ||
        eng_alloc(e, (void**)&e->S23, n3) || eng_alloc(e, (void**)&e->L0S1, n0) ||
        eng_alloc(e, (void**)&e->acc, n0 * 4) || eng_alloc(e, (void**)&e->macc, n0 * 4) ||
        eng_alloc(e, (void**)&e->s2snap, n3) || eng_alloc(e, (void**)&e->s2box, 64) ||
        eng_alloc(e, (void**)&e->T1p, n1 * 4) || eng_alloc(e, (void**)&e->T1c, n1 * 4) ||
        eng_alloc(e, (void**)&e->Tp2x, n2 * 4) || eng_alloc(e, (void**)&e->Tp3x, n3 * 4) ||
        eng_alloc(e, (void**)&e->preS2, n3) || eng_alloc(e, (void**)&e->mm, 64) ||
        eng_alloc(e, (void**)&e->reset3, n3) ||
        eng_alloc(e, (void**)&e->fa2, n2 * 4) || eng_alloc(e, (void**)&e->fb2, n2 * 4) ||
        eng_alloc(e, (void**)&e->fa3, n3 * 4) || eng_alloc(e, (void**)&e->fb3, n3 * 4) ||
        eng_alloc(e, (void**)&e->T3x, n3 * 4) ||
        eng_alloc(e, (void**)&e->d_desc, sizeof(GmStepDesc)) ||
        eng_alloc(e, (void**)&e->d_dm, sizeof(DevMat)) ||
        eng_alloc(e, (void**)&e->acc3, n3 * 4) || eng_alloc(e, (void**)&e->macc3, n3 * 4) ||
        eng_alloc(e, (void**)&e->src3,
                  (size_t)e->N23 *
                      (cfg->level[3].n[0] + cfg->level[3].n[1] + cfg->level[3].n[2]) * 4))
        return bad(-1);
    if (hipMemcpy(e->d_dm, &e->dm, sizeof(DevMat), hipMemcpyHostToDevice) != hipSuccess)
        return bad(fail("copy of the material constants failed"));
    for (int i = 0; i < 2; ++i) {
        if (eng_alloc(e, (void**)&e->T2b[i], n2 * 4) || eng_alloc(e, (void**)&e->S12b[i], n2) ||
            eng_alloc(e, (void**)&e->T3b[i], n3 * 4) || eng_alloc(e, (void**)&e->S13b[i], n3) ||
            eng_alloc(e, (void**)&e->S23b[i], n3))
            return bad(-1);
    }
    e->L2all.resize(e->N2);
    e->L3all.resize(e->N2);
    e->histTp.resize(e->N2);
    for (int i = 0; i < e->N2; ++i) {
        if (eng_alloc(e, (void**)&e->L2all[i], n2 * 4) ||
            eng_alloc(e, (void**)&e->L3all[i], n3 * 4) ||
            eng_alloc(e, (void**)&e->histTp[i], n3 * 4))
            return bad(-1);
    }
    {
        // k, cp companions of every L2 and L3 temperature buffer a
        // carried-property step reads or writes
        std::vector<std::pair<const float*, int64_t>> keys = {
            {e->T3, n3}, {e->T3b[0], n3}, {e->T3b[1], n3}, {e->T3x, n3},
            {e->T2, n2}, {e->T2b[0], n2}, {e->T2b[1], n2}};
        for (int i = 0; i < e->N2; ++i) {
            keys.push_back({e->L3all[i], n3});
            keys.push_back({e->L2all[i], n2});
        }
        for (const auto& k : keys) {
            float* b = nullptr;
            if (eng_alloc(e, (void**)&b, k.second * 8)) return bad(-1);
            e->kc_key.push_back(k.first);
            e->kc_T.push_back(b);
        }
    }
    // fine -> coarse pairs in the reference's Shapes order: L2->L1, L3->L1, L3->L2
    const int pc[3] = {1, 1, 2}, pf[3] = {2, 3, 3};
    for (int i = 0; i < 3; ++i) {
        if (pair_alloc(e->pair[i], pc[i], pf[i], cfg->level[pc[i]], cfg->level[pf[i]]))
            return bad(-1);
        if (eng_alloc(e, (void**)&e->part[i], pair_elems(e->pair[i].p) * 8 * sizeof(double)))
            return bad(-1);
        e->gmax[i] = pair_gmax(e->pair[i].p);
    }
    if (eng_alloc(e, (void**)&e->V1, pair_nodes(e->pair[0].p) * 4) ||
        eng_alloc(e, (void**)&e->V1m, pair_nodes(e->pair[0].p) * 4) ||
        eng_alloc(e, (void**)&e->V2, pair_nodes(e->pair[2].p) * 4) ||
        eng_alloc(e, (void**)&e->G1, (size_t)e->N23 * 3 * e->gmax[1] * 4) ||
        eng_alloc(e, (void**)&e->G2, (size_t)e->N23 * 3 * e->gmax[2] * 4) ||
        eng_alloc(e, (void**)&e->F2, (size_t)e->N2 * pair_nodes(e->pair[2].p) * 4) ||
        eng_alloc(e, (void**)&e->Fg, std::max(pair_nodes(e->pair[1].p),
                                              pair_nodes(e->pair[2].p)) * 4))
        return bad(-1);
    for (int i = 0; i < 4; ++i) {
        if (hipHostMalloc((void**)&e->ring[i], sizeof(GmStepDesc), hipHostMallocDefault) !=
            hipSuccess)
            return bad(fail("hipHostMalloc of the step descriptor failed"));
        if (hipEventCreateWithFlags(&e->ring_ev[i], hipEventDisableTiming) != hipSuccess)
            return bad(fail("hipEventCreate failed"));
        HIPCHK(hipEventRecord(e->ring_ev[i], e->s));
    }
    // initial state (SetupLevels, cF:150-157): T = T_amb, S1 = 0
    launch_fill_f32(e->T1, mat->T_amb, n1, e->s);
    launch_fill_f32(e->T2, mat->T_amb, n2, e->s);
    launch_fill_f32(e->T3, mat->T_amb, n3, e->s);
    HIPCHK(hipMemsetAsync(e->S11, 0, n1 * 4, e->s));
    HIPCHK(hipMemsetAsync(e->S1st, 0, (size_t)e->nst * n1 * 4, e->s));
    HIPCHK(hipMemsetAsync(e->Tp2, 0, n2 * 4, e->s));
    HIPCHK(hipMemsetAsync(e->Tp3, 0, n3 * 4, e->s));
    HIPCHK(hipMemsetAsync(e->S12, 0, n2, e->s));
    HIPCHK(hipMemsetAsync(e->S13, 0, n3, e->s));
    HIPCHK(hipMemsetAsync(e->S23, 0, n3, e->s));
    HIPCHK(hipMemsetAsync(e->reset3, 0, n3, e->s));
    HIPCHK(hipMemsetAsync(e->acc3, 0, n3 * 4, e->s));
    HIPCHK(hipMemsetAsync(e->macc3, 0, n3 * 4, e->s));
    HIPCHK(hipMemsetAsync(e->L0S1, 0, n0, e->s));
    HIPCHK(hipMemsetAsync(e->acc, 0, n0 * 4, e->s));
    HIPCHK(hipMemsetAsync(e->macc, 0, n0 * 4, e->s));
    HIPCHK(hipMemsetAsync(e->s2box, 0, 64, e->s));
    // descriptor: initial meshes and overlap boxes
    GmStepDesc& h = e->host;
    memset(&h, 0, sizeof(h));
    for (int l = 0; l < 4; ++l) h.g[l] = h.gprev[l] = cfg->level[l];
    for (int d = 0; d < 3; ++d) {
        h.ov3.lo[d] = cfg->ov_lo[3][d];
        h.ov3.n[d] = cfg->ov_n[3][d];
        h.ov3_xyz.n[d] = cfg->ov_n[3][d];
        h.ov3_xyz.x0[d] = cfg->ov_x0[3][d];
        h.ov3_xyz.h[d] = cfg->level[2].h[d];
        h.ov2.lo[d] = cfg->ov_lo[2][d];
        h.ov2.n[d] = cfg->ov_n[2][d];
        h.ov2_xyz.n[d] = cfg->ov_n[2][d];
        h.ov2_xyz.x0[d] = cfg->ov_x0[2][d];
        h.ov2_xyz.h[d] = cfg->level[1].h[d];
        h.l0_lo3[d] = cfg->ov_lo[0][d];
        h.l0_lo2[d] = cfg->ov_lo[1][d];
        h.l0_st2[d] = cfg->l0_stride_l2[d];
        h.l0_n[d] = cfg->level[0].n[d];
    }
    h.l1_ne_active = (int64_t)(cfg->level[1].n[0] - 1) * (cfg->level[1].n[1] - 1) *
                     (cfg->level[1].n[2] - 1);
    h.l1_nn_active = n1;
    HIPCHK(hipStreamSynchronize(e->s));
    if (upload_desc(e, desc_bytes(0))) return bad(-1);
    {
        const GmPair ps[3] = {e->pair[0].p, e->pair[1].p, e->pair[2].p};
        launch_axis_tables(ps, 3, e->d_desc, e->s);
    }
    for (int i = 0; i < 3; ++i)
        HIPCHK(hipMemsetAsync(e->part[i], 0, pair_elems(e->pair[i].p) * 8 * sizeof(double), e->s));
    e->fast_restrict = pairs_aligned(e);
    for (int i = 0; i < 3; ++i) strip_phase(e, i, e->strip_R0[i], e->strip_a0[i]);
    HIPCHK(hipStreamSynchronize(e->s));
    HIPCHK(hipGetLastError());
    *out = e;
    return 0;
}

extern "C" int gm_engine_field(gm_engine* e, int level, int field, void** ptr, int64_t* count) {
    void* p = nullptr;
    int64_t n = (level >= 0 && level < 4) ? e->nn[level] : 0;
    if (field >= GM_HIST_L2 && field <= GM_HIST_TP3) {
        if (level < 0 || level >= e->N2) return fail("history index out of range");
        p = field == GM_HIST_L2 ? (void*)e->L2all[level]
            : field == GM_HIST_L3 ? (void*)e->L3all[level] : (void*)e->histTp[level];
        *ptr = p;
        *count = field == GM_HIST_L2 ? e->nn[2] : e->nn[3];
        return 0;
    }
    switch (field) {
        case GM_T0: p = level == 1 ? (void*)e->T1 : level == 2 ? (void*)e->T2 : level == 3 ? (void*)e->T3 : nullptr; break;
        case GM_TPRIME0: p = level == 2 ? (void*)e->Tp2 : level == 3 ? (void*)e->Tp3 : nullptr; break;
        case GM_S1: p = level == 0 ? (void*)e->L0S1 : level == 1 ? (void*)e->S11 : level == 2 ? (void*)e->S12 : (void*)e->S13; break;
        case GM_S2: p = level == 3 ? (void*)e->S23 : nullptr; break;
        case GM_ACCUM: p = level == 0 ? (void*)e->acc : nullptr; break;
        case GM_MAX_ACCUM: p = level == 0 ? (void*)e->macc : nullptr; break;
        case GM_ACCUM_L3: p = e->acc3; n = e->nn[3]; break;
        case GM_MAX_ACCUM_L3: p = e->macc3; n = e->nn[3]; break;
        case GM_RESET_L3: p = e->reset3; n = e->nn[3]; break;
        default: break;
    }
    if (!p) return fail("no such field on this level");
    *ptr = p;
    *count = n;
    return 0;
}

extern "C" int gm_engine_grid(gm_engine* e, int level, gm_grid* out) {
    if (level < 0 || level > 3) return fail("level out of range");
    *out = e->host.g[level];
    return 0;
}

extern "C" int gm_engine_set_layer(gm_engine* e, int64_t ne_act, int64_t nn_act,
                                   const int64_t* sub4) {
    e->host.l1_ne_active = ne_act;
    e->host.l1_nn_active = nn_act;
    for (int i = 0; i < 4; ++i) e->host.sub[i] = sub4[i];
    e->graph_ok = false;  // static launch arguments changed
    return 0;
}

extern "C" int gm_engine_use_graph(gm_engine* e, int enable) {
    e->use_graph = enable != 0;
    e->graph_ok = false;
    return 0;
}

// ---- engine state for checkpoints (go_melt.py:115-129, 390-411) -------------

// The persistent device arrays of the engine except Level0 (whose S1 and
// melt-time maps the caller saves itself, as the reference's checkpoint
// holds Levels[0] and the two accum arrays), in a fixed order.
static std::vector<std::pair<void*, size_t>> state_arrays(gm_engine* e) {
    const size_t n1 = e->nn[1], n2 = e->nn[2], n3 = e->nn[3];
    return {{e->T1, n1 * 4},  {e->S11, n1 * 4}, {e->S1st, (size_t)e->nst * n1 * 4},
            {e->T2, n2 * 4},  {e->Tp2, n2 * 4}, {e->S12, n2},
            {e->T3, n3 * 4},  {e->Tp3, n3 * 4}, {e->S13, n3},
            {e->S23, n3},     {e->preS2, n3},   {e->reset3, n3},
            {e->s2snap, n3},  {e->s2box, 64},   {e->acc3, n3 * 4},
            {e->macc3, n3 * 4}};
}

static const uint32_t kStateMagic = 0x54534d47u;  // "GMST"

static size_t state_header_bytes() {
    return 8 + sizeof(gm_config) + sizeof(GmStepDesc) + 6 * sizeof(int);
}

extern "C" int gm_engine_state_bytes(gm_engine* e, int64_t* bytes) {
    size_t b = state_header_bytes();
    for (const auto& a : state_arrays(e)) b += a.second;
    *bytes = (int64_t)b;
    return 0;
}

extern "C" int gm_engine_save_state(gm_engine* e, void* host) {
    char* p = (char*)host;
    const uint32_t hdr[2] = {kStateMagic, (uint32_t)sizeof(GmStepDesc)};
    memcpy(p, hdr, 8);
    p += 8;
    memcpy(p, &e->cfg, sizeof(gm_config));
    p += sizeof(gm_config);
    memcpy(p, &e->host, sizeof(GmStepDesc));
    p += sizeof(GmStepDesc);
    memcpy(p, e->s3, 3 * sizeof(int));
    memcpy(p + 3 * sizeof(int), e->s2, 3 * sizeof(int));
    p += 6 * sizeof(int);
    for (const auto& a : state_arrays(e)) {
        HIPCHK(hipMemcpyAsync(p, a.first, a.second, hipMemcpyDeviceToHost, e->s));
        p += a.second;
    }
    HIPCHK(hipStreamSynchronize(e->s));
    return 0;
}

extern "C" int gm_engine_load_state(gm_engine* e, const void* host, int64_t bytes) {
    int64_t want = 0;
    gm_engine_state_bytes(e, &want);
    if (bytes != want) return fail("engine state: size does not match this deck");
    const char* p = (const char*)host;
    uint32_t hdr[2];
    memcpy(hdr, p, 8);
    if (hdr[0] != kStateMagic || hdr[1] != (uint32_t)sizeof(GmStepDesc))
        return fail("engine state: not a state of this library");
    p += 8;
    if (memcmp(p, &e->cfg, sizeof(gm_config)) != 0)
        return fail("engine state: saved from a different deck");
    p += sizeof(gm_config);
    memcpy(&e->host, p, sizeof(GmStepDesc));
    p += sizeof(GmStepDesc);
    memcpy(e->s3, p, 3 * sizeof(int));
    memcpy(e->s2, p + 3 * sizeof(int), 3 * sizeof(int));
    p += 6 * sizeof(int);
    HIPCHK(hipStreamSynchronize(e->s));
    for (const auto& a : state_arrays(e)) {
        HIPCHK(hipMemcpy(a.first, p, a.second, hipMemcpyHostToDevice));
        p += a.second;
    }
    // derived state: restriction path and strip layout, descriptor, tables
    e->graph_ok = false;
    e->fast_restrict = !pairs_aligned(e);  // force update_alignment to clear partials
    if (update_alignment(e)) return -1;
    if (upload_desc(e, desc_bytes(0))) return -1;
    {
        const GmPair ps[3] = {e->pair[0].p, e->pair[1].p, e->pair[2].p};
        launch_axis_tables(ps, 3, e->d_desc, e->s);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(e->s));
    return 0;
}

extern "C" int gm_engine_set_accum_mode(gm_engine* e, int external) {
    const bool x = external != 0;
    if (x != e->acc_external) {
        e->acc_external = x;
        e->graph_ok = false;  // the recorded step gathers/scatters Level0 or not
    }
    return 0;
}

extern "C" int gm_engine_overlap_box(gm_engine* e, int which, int32_t* lo, int32_t* n,
                                     int32_t* stride) {
    const GmStepDesc& h = e->host;
    const gm_config& c = e->cfg;
    for (int d = 0; d < 3; ++d) {
        switch (which) {
            case 0: lo[d] = h.l0_lo3[d]; n[d] = c.ov_n[0][d]; stride[d] = 1; break;
            case 1: lo[d] = h.l0_lo2[d]; n[d] = c.ov_n[1][d]; stride[d] = h.l0_st2[d]; break;
            case 2: lo[d] = h.ov2.lo[d]; n[d] = h.ov2.n[d]; stride[d] = 1; break;
            case 3: lo[d] = h.ov3.lo[d]; n[d] = h.ov3.n[d]; stride[d] = 1; break;
            default: return fail("overlap box: which must be 0..3");
        }
    }
    return 0;
}

// Re-checks the node alignment of the pairs after the meshes changed; a
// change of restriction path clears the partial buffers and the graph.
static int update_alignment(gm_engine* e) {
    const bool a = pairs_aligned(e);
    if (a != e->fast_restrict) {
        e->fast_restrict = a;
        e->graph_ok = false;
        for (int i = 0; i < 3; ++i)
            HIPCHK(hipMemsetAsync(e->part[i], 0, pair_elems(e->pair[i].p) * 8 * sizeof(double),
                                  e->s));
    }
    // the wave-brick restriction's strip layout (its grid size) follows the
    // pairs' x ratio and phase; a recorded graph holds the old grid
    for (int i = 0; i < 3; ++i) {
        int R0, a0;
        strip_phase(e, i, R0, a0);
        if (R0 != e->strip_R0[i] || a0 != e->strip_a0[i]) {
            e->strip_R0[i] = R0;
            e->strip_a0[i] = a0;
            e->graph_ok = false;
        }
    }
    return 0;
}

// ---- moveEverything (cF:2400-2510) ----------------------------------------

extern "C" int gm_engine_move(gm_engine* e, const float* v, const float* vstart,
                              int32_t* move_v_out) {
    const gm_config& c = e->cfg;
    GmStepDesc& h = e->host;
    float vtot[3];
    for (int d = 0; d < 3; ++d) vtot[d] = v[d] - vstart[d];
    float v3[3], v2[3];
    for (int d = 0; d < 3; ++d) {
        v3[d] = std::min(std::max(vtot[d], c.move_lo[1][d]), c.move_hi[1][d]);
        v2[d] = std::min(std::max(vtot[d], c.move_lo[0][d]), c.move_hi[0][d]);
    }
    for (int l = 0; l < 4; ++l) h.gprev[l] = h.g[l];
    const float h2[3] = {c.level[2].h[0], c.level[2].h[1], c.level[2].h[2]};
    const float h1[3] = {c.level[1].h[0], c.level[1].h[1], c.level[1].h[2]};
    const float hL1[3] = {h1[0], h1[1], c.layer_height};
    int s3[3], s2[3], mv[3];
    for (int d = 0; d < 3; ++d) {
        s3[d] = shift_of(v3[d], h2[d]);
        s2[d] = shift_of(v2[d], hL1[d]);
        mv[d] = s2[d] * c.ratio12[d];
    }
    for (int d = 0; d < 3; ++d) {
        h.g[3].x0[d] = c.level[3].x0[d] + h2[d] * (float)s3[d];
        h.g[2].x0[d] = c.level[2].x0[d] + hL1[d] * (float)s2[d];
    }
    // L3 overlap (L2 nodes inside L3), cF:1727-1764 then cF:2472-2474
    for (int d = 0; d < 3; ++d) {
        h.ov3.lo[d] = c.ov_lo[3][d] + s3[d] - mv[d];
        h.ov3_xyz.x0[d] = c.ov_x0[3][d] + h2[d] * (float)s3[d];
    }
    // L2 overlap (L1 nodes inside L2), cF:1767-1804
    for (int d = 0; d < 3; ++d) {
        const int sh = shift_of(v2[d], h1[d]);
        h.ov2.lo[d] = c.ov_lo[2][d] + sh;
        if (d < 2) h.ov2_xyz.x0[d] = c.ov_x0[2][d] + h1[d] * (float)sh;
    }
    h.ov2_xyz.x0[2] = c.ov_x0[2][2] + c.layer_height * (float)shift_of(v2[2], c.layer_height);
    // Level0 boxes, cF:2477-2497
    const float h0z[3] = {h1[0], h1[1], h2[2]};
    const int r2[3] = {c.ratio12[0] * c.ratio23[0], c.ratio12[1] * c.ratio23[1], c.ratio23[2]};
    for (int d = 0; d < 3; ++d) {
        h.l0_lo3[d] = c.ov_lo[0][d] + c.ratio23[d] * shift_of(v3[d], h2[d]);
        h.l0_lo2[d] = c.ov_lo[1][d] + r2[d] * shift_of(v2[d], h0z[d]);
    }
    h.l0_lo3[2] -= mv[2] * c.ratio23[2];
    h.l0_lo2[2] -= mv[2] * c.ratio23[2];
    for (int d = 0; d < 3; ++d) {
        e->s3[d] = s3[d];
        e->s2[d] = s2[d];
        if (move_v_out) move_v_out[d] = mv[d];
    }
    if (update_alignment(e)) return -1;
    if (upload_desc(e, desc_bytes(0))) return -1;
    const int64_t n3 = e->nn[3], n2 = e->nn[2];
    launch_move_l3(e->d_desc, e->T1, e->Tp2, e->Tp3, e->T3b[0], e->Tp3x, n3, e->s);
    launch_move_l2(e->d_desc, e->T1, e->Tp2, e->T2b[0], e->Tp2x, n2, e->s);
    {
        CopyList L;
        L.c[0] = CopyItem{e->T3, e->T3b[0], n3 * 4};
        L.c[1] = CopyItem{e->Tp3, e->Tp3x, n3 * 4};
        L.c[2] = CopyItem{e->T2, e->T2b[0], n2 * 4};
        L.c[3] = CopyItem{e->Tp2, e->Tp2x, n2 * 4};
        launch_copy_multi(L, 4, e->s);
    }
    {
        const GmPair ps[3] = {e->pair[0].p, e->pair[1].p, e->pair[2].p};
        launch_axis_tables(ps, 3, e->d_desc, e->s);
    }
    launch_l0_gather(e->d_desc, e->L0S1, e->s2snap, e->s2box, e->S13, e->S23, e->S12, n3, n2,
                     e->s);
    HIPCHK(hipGetLastError());
    return 0;
}

// ---- level-step helpers -----------------------------------------------------

// L1 solve (computeL1Temperature cF:2813-2854 + clamp)
static void run_l1(gm_engine* e, const float* corr, float* Tout, int dt_mode, float dt,
                   int clamp) {
    LSArgs a = ls_default();
    a.level = 1;
    a.T = e->T1;
    a.S1 = e->S11;
    a.s1_float = 1;
    a.Tout = Tout;
    a.n_sub = e->host.sub[1];
    a.ne_active = e->host.l1_ne_active;
    a.nn_active = e->host.l1_nn_active;
    a.dt_mode = dt_mode;
    a.dt = dt;
    a.conv = 1;
    if (corr) {
        a.corr = corr;
        for (int d = 0; d < 3; ++d) {
            a.corr_info[d] = e->pair[0].p.ax[d].info;
            a.corr_nmax[d] = e->pair[0].p.cmax[d] + 1;
        }
    }
    a.bc_mode = GM_BC_L1;
    for (int i = 0; i < 5; ++i) a.bc_vals[i] = e->cfg.l1_bc[i];
    a.clamp = clamp;
    memcpy(a.kc, e->kc[1], sizeof(a.kc));
    a.mrow = e->mrow[1];
    a.w2 = e->w2[1];
    a.w3 = e->w3[1];
    a.prec_ok = e->dm.pnb > 0;
    launch_level_step(a, e->d_desc, e->cfg.level[1], e->d_dm, e->s);
}

// L2 solve (computeL2Temperature cF:2917-2957 + clamp)
static void run_l2(gm_engine* e, const float* Tin, const uint8_t* S1in, uint8_t* S1out,
                   const float* corr, float alpha, float* Tout, int dt_mode, int dt_index,
                   float dt, int clamp) {
    LSArgs a = ls_default();
    a.level = 2;
    a.T = Tin;
    a.S1 = S1in;
    a.S1out = S1out;
    a.Tout = Tout;
    a.n_sub = e->host.sub[2];
    a.ne_active = (int64_t)(e->cfg.level[2].n[0] - 1) * (e->cfg.level[2].n[1] - 1) *
                  (e->cfg.level[2].n[2] - 1);
    a.nn_active = e->nn[2];
    a.dt_mode = dt_mode;
    a.dt_index = dt_index;
    a.dt = dt;
    a.conv = 1;
    a.corr = corr;
    for (int d = 0; d < 3; ++d) {
        a.corr_info[d] = e->pair[2].p.ax[d].info;
        a.corr_nmax[d] = e->pair[2].p.cmax[d] + 1;
    }
    a.bc_mode = GM_BC_FINE;
    a.bc_level = 1;
    a.bcA = e->fa2;
    a.bcB = e->fb2;
    a.alpha = alpha;
    a.clamp = clamp;
    memcpy(a.kc, e->kc[2], sizeof(a.kc));
    a.mrow = e->mrow[2];
    a.w2 = e->w2[2];
    a.w3 = e->w3[2];
    launch_level_step(a, e->d_desc, e->cfg.level[2], e->d_dm, e->s);
}

// L3 solve (computeSolutions_L3 cF:3015-3054 + clamp [+ melt time])
static void run_l3(gm_engine* e, const float* Tin, const uint8_t* S1in, uint8_t* S1out, int row,
                   float alpha, float* Tout, int clamp, const uint8_t* S2prev, uint8_t* S2out) {
    LSArgs a = ls_default();
    a.level = 3;
    a.T = Tin;
    a.S1 = S1in;
    a.S1out = S1out;
    a.Tout = Tout;
    a.n_sub = e->host.sub[3];
    a.ne_active = (int64_t)(e->cfg.level[3].n[0] - 1) * (e->cfg.level[3].n[1] - 1) *
                  (e->cfg.level[3].n[2] - 1);
    a.nn_active = e->nn[3];
    a.dt_mode = GM_DT_ROW;
    a.dt_index = row;
    a.src_mode = GM_SRC_L3;
    a.src_row = row;
    a.conv = 1;
    a.bc_mode = GM_BC_FINE;
    a.bc_level = 2;
    a.bcA = e->fa3;
    a.bcB = e->fb3;
    a.alpha = alpha;
    a.clamp = clamp;
    if (S2out) {
        a.accum_on = 1;
        a.S2prev = S2prev;
        a.S2out = S2out;
        a.accum = e->acc;
        a.max_accum = e->macc;
    }
    memcpy(a.kc, e->kc[3], sizeof(a.kc));
    a.mrow = e->mrow[3];
    a.w2 = e->w2[3];
    a.w3 = e->w3[3];
    launch_level_step(a, e->d_desc, e->cfg.level[3], e->d_dm, e->s);
}

// k, cp companion of an L3 temperature buffer
static float* kc_of(gm_engine* e, const float* T) {
    for (size_t i = 0; i < e->kc_key.size(); ++i)
        if (e->kc_key[i] == T) return e->kc_T[i];
    return nullptr;
}

// L2 solve of the subcycle with the nodal properties carried between steps
static LSArgs l2_args_kc(gm_engine* e, const float* Tin, const uint8_t* S1in, uint8_t* S1out,
                         const float* corr, float alpha, float* Tout, int dt_index,
                         const float* F = nullptr, float gscale = 0.0f) {
    LSArgs a = ls_default();
    a.level = 2;
    a.T = Tin;
    a.S1 = S1in;
    a.S1out = S1out;
    a.Tout = Tout;
    a.kcin = kc_of(e, Tin);
    a.kcout = kc_of(e, Tout);
    a.n_sub = e->host.sub[2];
    a.dt_mode = GM_DT_L2;
    a.dt_index = dt_index;
    a.conv = 1;
    if (corr) {
        a.corr = corr;
    } else {
        // restricted terms straight from the pair's FP64 node box, plus the
        // L2 source of the substep's positions (k_restrict_nodes_acc's sum)
        a.corr64 = e->part[2];
        a.corrG = F;
        a.corrG_scale = gscale;
    }
    for (int d = 0; d < 3; ++d) {
        a.corr_info[d] = e->pair[2].p.ax[d].info;
        a.corr_nmax[d] = e->pair[2].p.cmax[d] + 1;
    }
    a.bcA = e->fa2;
    a.bcB = e->fb2;
    a.alpha = alpha;
    a.clamp = 1;
    memcpy(a.kc, e->kc[2], sizeof(a.kc));
    a.mrow = e->mrow[2];
    a.w2 = e->w2[2];
    a.w3 = e->w3[2];
    a.prec_ok = e->dm.pnb > 0;
    return a;
}

static void run_l2_kc(gm_engine* e, const float* Tin, const uint8_t* S1in, uint8_t* S1out,
                      const float* corr, float alpha, float* Tout, int dt_index,
                      const float* F = nullptr, float gscale = 0.0f) {
    const LSArgs a = l2_args_kc(e, Tin, S1in, S1out, corr, alpha, Tout, dt_index, F, gscale);
    launch_level_step_kc(a, e->d_desc, e->cfg.level[2], e->d_dm, e->s);
}

// L3 solve of the subcycle with the nodal properties carried between steps
static LSArgs l3_args_kc(gm_engine* e, const float* Tin, const uint8_t* S1in, uint8_t* S1out,
                         int row, float alpha, float* Tout, const uint8_t* S2prev,
                         uint8_t* S2out) {
    LSArgs a = ls_default();
    a.level = 3;
    a.T = Tin;
    a.S1 = S1in;
    a.S1out = S1out;
    a.Tout = Tout;
    a.kcin = kc_of(e, Tin);
    a.kcout = kc_of(e, Tout);
    a.srcT = e->src3;
    a.n_sub = e->host.sub[3];
    a.dt_mode = GM_DT_ROW;
    a.dt_index = row;
    a.src_mode = GM_SRC_L3;
    a.src_row = row;
    a.conv = 1;
    a.bcA = e->fa3;
    a.bcB = e->fb3;
    a.alpha = alpha;
    a.clamp = 1;
    if (S2out) {
        a.accum_on = 1;
        a.S2prev = S2prev;
        a.S2out = S2out;
        a.accum = e->acc3;  // per L3 node (gathered from Level0 for Part 2)
        a.max_accum = e->macc3;
    }
    memcpy(a.kc, e->kc[3], sizeof(a.kc));
    a.mrow = e->mrow[3];
    a.w2 = e->w2[3];
    a.w3 = e->w3[3];
    a.prec_ok = e->dm.pnb > 0;
    return a;
}

static void run_l3_kc(gm_engine* e, const float* Tin, const uint8_t* S1in, uint8_t* S1out,
                      int row, float alpha, float* Tout, const uint8_t* S2prev, uint8_t* S2out) {
    const LSArgs a = l3_args_kc(e, Tin, S1in, S1out, row, alpha, Tout, S2prev, S2out);
    launch_level_step_kc(a, e->d_desc, e->cfg.level[3], e->d_dm, e->s);
}

static RArgs rargs(gm_engine* e, int fine, const float* T, const uint8_t* S1, const float* Tp,
                   const float* TpA, const float* TpB, int dt_mode, int dt_index, float dt) {
    RArgs r;
    memset(&r, 0, sizeof(r));
    r.T = T;
    r.S1 = S1;
    r.n_sub = e->host.sub[fine];
    r.Tp = Tp;
    r.TpA = TpA;
    r.TpB = TpB;
    r.dt_mode = dt_mode;
    r.dt_index = dt_index;
    r.dt = dt;
    r.wf = e->w3[fine];
    return r;
}

// Fine nodes on every coarse node of the overlap along each axis: integer
// spacing ratio, fine origin an integer number of fine spacings from the
// coarse origin, fine extent inside the coarse mesh.
static bool grids_aligned(const gm_grid& gc, const gm_grid& gf) {
    for (int d = 0; d < 3; ++d) {
        const double R = (double)gc.h[d] / gf.h[d];
        const double off = ((double)gf.x0[d] - gc.x0[d]) / gf.h[d];
        if (std::fabs(R - std::round(R)) > 1e-3 || std::round(R) < 1.0) return false;
        if (std::fabs(off - std::round(off)) > 1e-3 || std::round(off) < 0.0) return false;
        if (std::round(off) + (gf.n[d] - 1) > std::round(R) * (gc.n[d] - 1) + 1e-6) return false;
    }
    return true;
}

// x spacing ratio of pair i and the first fine x index on a coarse node
// (R0 = 0 when the pair is not node-aligned)
static void strip_phase(gm_engine* e, int i, int& R0, int& a0) {
    const gm_grid& gc = e->host.g[e->pair[i].p.coarse];
    const gm_grid& gf = e->host.g[e->pair[i].p.fine];
    R0 = 0;
    a0 = 0;
    if (!grids_aligned(gc, gf)) return;
    const int R = (int)std::lround((double)gc.h[0] / gf.h[0]);
    const int off = (int)std::lround(((double)gf.x0[0] - gc.x0[0]) / gf.h[0]);
    if (R < 1 || R > 61 || off < 0) return;
    R0 = R;
    a0 = (R - off % R) % R;
}

static bool pairs_aligned(gm_engine* e) {
    bool ok = true;
    for (int i = 0; i < 3; ++i)
        ok = ok && grids_aligned(e->host.g[e->pair[i].p.coarse], e->host.g[e->pair[i].p.fine]);
    return ok;
}

// Restricted T' terms of pair i into its partial buffer: node-centric FP64
// node box when the pairs are aligned, element partials otherwise.
// the node-aligned restriction's arguments of pair i (fast_restrict engines)
static RArgs fast_rargs(gm_engine* e, int i, RArgs r) {
    const GmPair& p = e->pair[i].p;
    if (e->use_kc && p.fine >= 2) r.KC = kc_of(e, r.T);
    r.R0 = e->strip_R0[i];
    r.a0 = e->strip_a0[i];
    r.dpp_ok = r.R0 > 0;
    if (r.dpp_ok) {
        const gm_grid& gc = e->host.g[p.coarse];
        const gm_grid& gf = e->host.g[p.fine];
        r.R1 = (int)std::lround((double)gc.h[1] / gf.h[1]);
        r.R2 = (int)std::lround((double)gc.h[2] / gf.h[2]);
        r.W = r.R0 * (61 / r.R0);
        r.invR[0] = 1.0 / r.R0;
        r.invR[1] = 1.0 / r.R1;
        r.invR[2] = 1.0 / r.R2;
        r.dpp_ok = r.R1 >= 1 && r.R2 >= 1;
    }
    memcpy(r.kcs, e->kc[p.fine], sizeof(r.kcs));
    memcpy(r.mcs, e->mcs[p.fine], sizeof(r.mcs));
    return r;
}

static void eng_restrict(gm_engine* e, int i, RArgs r) {
    const GmPair& p = e->pair[i].p;
    if (e->fast_restrict)
        launch_restrict_fast(p, e->d_desc, e->d_dm, fast_rargs(e, i, r), e->part[i], e->s);
    else
        launch_restrict(p, e->d_desc, e->dm, r, e->part[i], e->s);
}

// eng_restrict of pair i as an item of the batched restriction, when the
// pair takes the form the batched launch runs
static bool restrict_item(gm_engine* e, int i, const RArgs& r0, RItem* it) {
    if (!e->fast_restrict) return false;
    const RArgs r = fast_rargs(e, i, r0);
    if (!restrict_batchable(r)) return false;
    it->p = e->pair[i].p;
    it->d = e->d_desc;
    it->r = r;
    it->acc = e->part[i];
    return true;
}

static void eng_restrict_nodes(gm_engine* e, int outer, int inner, const float* Vbase, float* V,
                               int gpair, const float* G, int npos, int nmax, float scale) {
    const GmPair* pi = inner >= 0 ? &e->pair[inner].p : nullptr;
    const GmPair* pg = gpair >= 0 ? &e->pair[gpair].p : nullptr;
    if (e->fast_restrict) {
        if (pg && G) launch_src_box(*pg, G, 1, npos, nmax, e->Fg, e->s);
        launch_restrict_nodes_acc(e->pair[outer].p, e->part[outer], pi,
                                  inner >= 0 ? e->part[inner] : nullptr, Vbase, V, pg,
                                  (pg && G) ? e->Fg : nullptr, scale, e->s);
    } else
        launch_restrict_nodes(e->pair[outer].p, e->part[outer], pi,
                              inner >= 0 ? e->part[inner] : nullptr, Vbase, V, pg, G, npos, nmax,
                              scale, e->s);
}

// getNewTprime (cF:2060-2099) for fine level f in {2, 3}.
// cS1: S1 state of the coarse field, to refresh its carried k, cp at the
// injected nodes (L2 coarse fields of the subcycle).
static void new_tprime(gm_engine* e, int f, const float* fineT, float* coarseT, float* Tp,
                       const uint8_t* cS1 = nullptr) {
    const GmBox& b = f == 3 ? e->host.ov3 : e->host.ov2;
    const int64_t nb = (int64_t)b.n[0] * b.n[1] * b.n[2];
    float* cKC = (e->use_kc && cS1) ? kc_of(e, coarseT) : nullptr;
    if (e->fast_restrict) {
        launch_tprime_inject(e->d_desc, f, f - 1, fineT, coarseT, Tp, e->cfg.level[f], nb,
                             e->d_dm, cS1, e->host.sub[f - 1], cKC, e->s);
        return;
    }
    launch_inject(e->d_desc, f, f - 1, fineT, coarseT, nb, e->s);
    launch_tprime(e->d_desc, f, f - 1, fineT, coarseT, Tp, e->nn[f], e->s);
    if (cKC) launch_props_kc(e->d_dm, coarseT, cS1, e->nn[f - 1], e->host.sub[f - 1], cKC, e->s);
}

static void l1_s1_update(gm_engine* e) {
    const GmBox& b = e->host.ov2;
    launch_l1_s1_from_l2(e->d_desc, e->S12, e->S11, (int64_t)b.n[0] * b.n[1] * b.n[2], e->s);
    launch_fill_f32(e->S11, 1.0f, e->host.sub[1], e->s);
}

// ---- subcycleGOMELT (cF:3224-3632) -------------------------------------------

// L2 solve of substep k with its restricted terms and source: read straight
// from the pair's FP64 node box by the step (aligned levels), or finished
// into V2 first
static void l2_substep(gm_engine* e, const float* T2cur, const uint8_t* S12cur, uint8_t* S12new,
                       const float* Gk, float alpha, float* T2new, int k) {
    const float gscale = e->w3[3] / (float)e->N3;
    if (e->fast_restrict) {
        // source box of substep k, from launch_src_box at the start of the step
        run_l2_kc(e, T2cur, S12cur, S12new, nullptr, alpha, T2new, k,
                  e->F2 + (size_t)k * pair_nodes(e->pair[2].p), gscale);
        return;
    }
    eng_restrict_nodes(e, 2, -1, nullptr, e->V2, 2, Gk, e->N3, e->gmax[2], gscale);
    run_l2_kc(e, T2cur, S12cur, S12new, e->V2, alpha, T2new, k);
}

// The subcycled L1 step runs as a fixed sequence of stages: the prologue
// (sources, L1 predictor), per L2 substep k of Part 1 the L2 stage
// (restriction, L2 solve, L3 face values) and the L3 stage (N3 L3 solves,
// then T' and injection), the middle (L1 corrector), Part 2 likewise, then the
// epilogue.  subcycle_body runs them all (the recorded graph); the staged
// entry points let a caller replace the N3 L3 solves of a substep by its own
// update of the L3 window (the learned operator).

static void stage_reset(gm_engine* e, SubcycleState& st) {
    st.T2cur = e->T2;
    st.S12cur = e->S12;
    st.T3cur = e->T3;
    st.Tp3cur = e->Tp3;
    st.S13cur = e->S13;
    st.S23cur = e->S23;
    st.b2 = st.b3 = 0;
    st.T2new = nullptr;
    st.S12new = nullptr;
    st.T3out = nullptr;
    st.T3ext = nullptr;
}

static void stage_prologue(gm_engine* e, SubcycleState& st) {
    const int N2 = e->N2, N3 = e->N3, N23 = e->N23;
    const int64_t n2 = e->nn[2], n3 = e->nn[3];
    GmStepDesc* d = e->d_desc;
    hipStream_t s = e->s;
    // L1 S1 from the L2 overlap, then substrate (cF:3273-3278)
    l1_s1_update(e);
    // k, cp of the start-of-step L3 state for the carried-property L3 steps
    launch_props_kc(e->d_dm, e->T3, e->S13, n3, e->host.sub[3], kc_of(e, e->T3), s, e->T2,
                    e->S12, n2, e->host.sub[2], kc_of(e, e->T2));
    // separable L3 source factors of all N2*N3 rows (the L3 mesh is fixed
    // during the step)
    {
        const gm_grid& g3 = e->cfg.level[3];
        launch_src3_tables(d, e->d_dm, N23, g3.n[0] + g3.n[1] + g3.n[2], e->w3[3], e->src3, s);
    }
    e->use_kc = true;
    // L1 source over all N2*N3 positions (cF:3285-3287)
    launch_source_tables(e->pair[1].p, d, e->dm, 0, N23, 0, nullptr, e->G1, e->gmax[1], s);
    // L2 sources per L2 substep (cF:3329-3331), shared by both passes: the
    // tables of substep k are positions k*N3 .. k*N3+N3-1, one launch for all
    launch_source_tables(e->pair[2].p, d, e->dm, 0, N23, 0, nullptr, e->G2, e->gmax[2], s);
    if (e->fast_restrict) launch_src_box(e->pair[2].p, e->G2, N2, N3, e->gmax[2], e->F2, s);
    // L1 flux terms from L3 and L2 T' (computeL1TprimeTerms_Part1, cF:2733-2810)
    eng_restrict(e, 1,
                    rargs(e, 3, e->T3, e->S13, e->Tp3, nullptr, nullptr, GM_DT_VALUE, 0, 1.0f));
    eng_restrict(e, 0,
                    rargs(e, 2, e->T2, e->S12, e->Tp2, nullptr, nullptr, GM_DT_VALUE, 0, 1.0f));
    eng_restrict_nodes(e, 0, 1, nullptr, e->V1,
                          1, e->G1, N23, e->gmax[1], e->w3[3] / (float)N23);
    // L1 predictor; its L2 face values for the Part-1 L2 substeps
    run_l1(e, e->V1, e->T1p, GM_DT_L1, 0.0f, 1);
    launch_face_bc(d, 2, 1, e->cfg.level[2], e->T1p, e->T1, e->fa2, e->fb2, s);
    stage_reset(e, st);
}

// The L2 stage before the L2 solve: the arguments of the restriction of the
// L3 T' terms onto L2, and the buffers of the substep's new L2 field.
static RArgs stage_l2_rargs(gm_engine* e, SubcycleState& st, int part, int k) {
    RArgs r;
    if (part == 1) {
        r = rargs(e, 3, st.T3cur, st.S13cur, st.Tp3cur, nullptr, nullptr, GM_DT_VALUE, 0, 1.0f);
        st.T2new = e->T2b[st.b2];
    } else {
        // flux + mass restriction onto L2 (cF:3496-3508)
        r = rargs(e, 3, st.T3cur, st.S13cur, st.Tp3cur, e->histTp[k], st.Tp3cur, GM_DT_L2, k,
                  0.0f);
        st.T2new = e->L2all[k];
    }
    st.S12new = e->S12b[st.b2];
    st.b2 ^= 1;
    return r;
}

static void stage_l2_pre(gm_engine* e, SubcycleState& st, int part, int k) {
    eng_restrict(e, 2, stage_l2_rargs(e, st, part, k));
}

// after the L2 solve and the L3 face values: the buffer that will hold the
// substep's final L3 field; a caller's own field goes there too unless that
// is the start field's buffer (Part 1 with an even N3 ends its solves in the
// buffer it started from), which must stay intact for a fallback to the
// plain solves
static void stage_l2_post(gm_engine* e, SubcycleState& st, int part, int k) {
    if (part == 1) st.T3out = e->T3b[(st.b3 + e->N3 - 1) & 1];
    else st.T3out = e->L3all[k];
    st.T3ext = st.T3out == st.T3cur ? e->T3x : st.T3out;
}

// Part 1 (part = 1, cF:3333-3420) or Part 2 (part = 2, cF:3466-3608) L2 stage
// of substep k: restriction of the L3 T' terms, L2 solve, L3 face values.
static void stage_l2(gm_engine* e, SubcycleState& st, int part, int k) {
    const float fN2 = (float)e->N2;
    const float* Gk = e->G2 + (size_t)k * e->N3 * 3 * e->gmax[2];
    stage_l2_pre(e, st, part, k);
    l2_substep(e, st.T2cur, st.S12cur, st.S12new, Gk, (float)(k + 1) / fN2, st.T2new, k);
    launch_face_bc(e->d_desc, 3, 2, e->cfg.level[3], st.T2new, st.T2cur, e->fa3, e->fb3, e->s);
    stage_l2_post(e, st, part, k);
}

// the L2 solve of stage_l2 as arguments of the batched L2 step (the
// node-aligned path, l2_substep's fast branch), dt by value
static LSArgs stage_l2_args(gm_engine* e, const SubcycleState& st, int k) {
    const float gscale = e->w3[3] / (float)e->N3;
    LSArgs a = l2_args_kc(e, st.T2cur, st.S12cur, st.S12new, nullptr,
                          (float)(k + 1) / (float)e->N2, st.T2new, k,
                          e->F2 + (size_t)k * pair_nodes(e->pair[2].p), gscale);
    a.grid = e->cfg.level[2];
    a.dt_mode = GM_DT_VALUE;
    a.dt = e->host.dt2[k];
    return a;
}

// L3 stage, closing: T' and injection into the new L2 field, then the
// substep's state becomes the next one's input.
static void stage_l3_close_state(gm_engine* e, SubcycleState& st, int part, int k,
                                 const uint8_t* S13fin, const uint8_t* S23fin, float* T3fin) {
    st.T2cur = st.T2new;
    st.S12cur = st.S12new;
    st.T3cur = T3fin;
    st.T3last = T3fin;
    st.S13cur = S13fin;
    if (part == 2) st.S23cur = S23fin;
    st.Tp3cur = e->histTp[k];
}

static void stage_l3_close(gm_engine* e, SubcycleState& st, int part, int k,
                           const uint8_t* S13fin, const uint8_t* S23fin, float* T3fin) {
    new_tprime(e, 3, T3fin, st.T2new, e->histTp[k], st.S12new);
    stage_l3_close_state(e, st, part, k, S13fin, S23fin, T3fin);
}

// new_tprime's node-aligned launch as an item of the batched T' + injection
static TpItem stage_tprime_item(gm_engine* e, const SubcycleState& st, int k, float* T3fin,
                                int* nblocks) {
    TpItem t;
    t.d = e->d_desc;
    t.fT = T3fin;
    t.cT = st.T2new;
    t.Tp = e->histTp[k];
    t.m = e->d_dm;
    t.cS1 = st.S12new;
    t.c_nsub = e->host.sub[2];
    t.cKC = e->use_kc ? kc_of(e, st.T2new) : nullptr;
    const GmBox& b = e->host.ov3;
    *nblocks = tprime_blocks(e->cfg.level[3], (int64_t)b.n[0] * b.n[1] * b.n[2], &t.ntp);
    return t;
}

// Buffers of the N3 L3 solves of substep k, solve by solve: Part 1 ping-pongs
// the field between the two work buffers and keeps S1 in a work buffer (the
// first solve writes all of it, later ones update it in place, written where
// it flips); Part 2 leaves its last field in the substep's history buffer and
// updates S1 and the melt state in place.
struct L3Walk {
    const float* T3in;
    const uint8_t* S13in;
    const uint8_t* S23in;
};
struct L3StepBufs {
    const float* Tin;
    const uint8_t* S1in;
    uint8_t* S1out;
    float* Tout;
    const uint8_t* S2in;
    uint8_t* S2out;
};

static L3Walk l3_walk_begin(const SubcycleState& st) { return {st.T3cur, st.S13cur, st.S23cur}; }

static L3StepBufs l3_walk_step(gm_engine* e, SubcycleState& st, L3Walk& w, int part, int k,
                               int j) {
    L3StepBufs b;
    b.Tin = w.T3in;
    b.S1in = w.S13in;
    if (part == 1) {
        b.Tout = e->T3b[st.b3];
        st.b3 ^= 1;
        b.S1out = e->S13b[0];
        b.S2in = nullptr;
        b.S2out = nullptr;
    } else {
        b.Tout = (j == e->N3 - 1) ? e->L3all[k] : e->T3b[st.b3];
        if (j != e->N3 - 1) st.b3 ^= 1;
        b.S1out = e->S13;
        b.S2in = w.S23in;
        b.S2out = e->S23;
    }
    w.T3in = b.Tout;
    w.S13in = b.S1out;
    if (part == 2) w.S23in = b.S2out;
    return b;
}

// L3 stage of substep k: the N3 L3 solves (computeSolutions_L3 cF:3015-3054
// with the time-blended faces), then T'.
static void stage_l3_plain(gm_engine* e, SubcycleState& st, int part, int k) {
    const int N3 = e->N3;
    const float fN3 = (float)N3;
    L3Walk w = l3_walk_begin(st);
    for (int j = 0; j < N3; ++j) {
        const L3StepBufs b = l3_walk_step(e, st, w, part, k, j);
        run_l3_kc(e, b.Tin, b.S1in, b.S1out, k * N3 + j, (float)(j + 1) / fN3, b.Tout, b.S2in,
                  b.S2out);
    }
    stage_l3_close(e, st, part, k, w.S13in, w.S23in, st.T3out);
}

// Phase flags and melt time of one jump of the L3 window over the N3 steps of
// L2 substep k (a caller's update T0 -> T1 in place of the N3 solves):
//   S1 = S1in | (T0 >= T_liq) | (T1 >= T_liq), substrate nodes solid;
//   with S2 (Part 2): a node is molten over the jump if T0 or T1 >= T_liq;
//   one that melts anew hands its accumulated time to max_accum and
//   restarts, as the plain step's per-step rule does, then adds the
//   substep's time dt2[k].
__global__ void k_jump_state(const GmStepDesc* __restrict__ d, const DevMat* __restrict__ m,
                             int k, const float* __restrict__ T0, const float* __restrict__ T1,
                             const uint8_t* __restrict__ S1in, uint8_t* __restrict__ S1out,
                             int64_t nsub, uint8_t* __restrict__ S2, float* __restrict__ acc,
                             float* __restrict__ macc, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float Tl = m->T_liq;
    const bool hot = T0[t] >= Tl || T1[t] >= Tl;
    S1out[t] = (uint8_t)((S1in[t] != 0 || hot || t < nsub) ? 1 : 0);
    if (S2) {
        const bool s2p = S2[t] != 0;
        if (hot) {
            const float a = acc[t], mx = macc[t];
            const float rr = s2p ? 0.0f : a;
            macc[t] = rr > mx ? rr : mx;
            acc[t] = a + d->dt2[k] - rr;
        }
        S2[t] = (uint8_t)hot;
    }
}

static void launch_jump_state(const GmStepDesc* d, const DevMat* m, int k, const float* T0,
                              const float* T1, const uint8_t* S1in, uint8_t* S1out, int64_t nsub,
                              uint8_t* S2, float* acc, float* macc, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_jump_state, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d, m, k,
                       T0, T1, S1in, S1out, nsub, S2, acc, macc, n);
}

// L3 stage of substep k when the caller has written the substep's final L3
// field into st.T3ext (one jump over the N3 steps): phase flags, melt time
// and the carried k, cp of that field, then T'.
static void stage_l3_external(gm_engine* e, SubcycleState& st, int part, int k) {
    const int64_t n3 = e->nn[3];
    float* T3fin = st.T3ext;
    uint8_t* S13o = part == 1 ? e->S13b[0] : e->S13;
    launch_jump_state(e->d_desc, e->d_dm, k, st.T3cur, T3fin, st.S13cur, S13o,
                      e->host.sub[3], part == 2 ? e->S23 : nullptr,
                      part == 2 ? e->acc3 : nullptr, part == 2 ? e->macc3 : nullptr, n3, e->s);
    launch_props_kc(e->d_dm, T3fin, S13o, n3, e->host.sub[3], kc_of(e, T3fin), e->s,
                    nullptr, nullptr, 0, 0, nullptr);
    // the ping-pong parity the N3 solves would leave (the next solves start in
    // the other buffer than the one holding the field; with T3x both are free)
    if (part == 1) st.b3 ^= (e->N3 & 1);
    else st.b3 ^= ((e->N3 - 1) & 1);
    stage_l3_close(e, st, part, k, S13o, part == 2 ? e->S23 : nullptr, T3fin);
}

static void stage_mid(gm_engine* e, SubcycleState& st) {
    GmStepDesc* d = e->d_desc;
    hipStream_t s = e->s;
    // L2 T' and L1 injection, then the L1 mass terms and corrector (cF:3433-3456)
    new_tprime(e, 2, st.T2cur, e->T1p, e->Tp2x);
    eng_restrict(e, 1,
                    rargs(e, 3, e->T3, e->S13, nullptr, st.Tp3cur, e->Tp3, GM_DT_L1, 0, 0.0f));
    eng_restrict(e, 0,
                    rargs(e, 2, e->T2, e->S12, nullptr, e->Tp2x, e->Tp2, GM_DT_L1, 0, 0.0f));
    eng_restrict_nodes(e, 0, 1, e->V1, e->V1m,
                          -1, nullptr, 0, 0, 0.0f);
    run_l1(e, e->V1m, e->T1c, GM_DT_L1, 0.0f, 1);
    launch_face_bc(d, 2, 1, e->cfg.level[2], e->T1c, e->T1, e->fa2, e->fb2, s);
    // ---- Part 2: restart L2/L3 from the start-of-step state ------------------
    // melt time of the L3 nodes: gathered from Level0 unless the caller
    // passed the window arrays (the driver's max_accum_time[idx], cF:3224-3238)
    if (!e->acc_external) launch_acc_l3(d, e->acc, e->macc, e->acc3, e->macc3, e->nn[3], 1, s);
    stage_reset(e, st);
}

static void stage_epilogue(gm_engine* e, SubcycleState& st) {
    const int64_t n1 = e->nn[1], n2 = e->nn[2], n3 = e->nn[3];
    GmStepDesc* d = e->d_desc;
    hipStream_t s = e->s;
    if (!e->acc_external) launch_acc_l3(d, e->acc, e->macc, e->acc3, e->macc3, n3, 0, s);
    // final L2 T' and L1 T0 (cF:3625-3627), then commit the end-of-step state
    new_tprime(e, 2, st.T2cur, e->T1c, e->Tp2);
    {
        CopyList L;
        L.c[0] = CopyItem{e->T2, st.T2cur, n2 * 4};
        L.c[1] = CopyItem{e->S12, st.S12cur, n2};
        L.c[2] = CopyItem{e->T3, st.T3cur, n3 * 4};
        L.c[3] = CopyItem{e->Tp3, st.Tp3cur, n3 * 4};
        L.c[4] = CopyItem{e->T1, e->T1c, n1 * 4};
        launch_copy_multi(L, 5, s);
    }
    // Level0 S1/S2 (cF:3628-3630)
    launch_l0_scatter(d, e->S13, e->S23, e->L0S1, e->s2snap, e->s2box, n3, s);
    e->use_kc = false;
}

static void subcycle_body(gm_engine* e) {
    SubcycleState st;
    stage_prologue(e, st);
    // ---- Part 1: L2 x N2 { L3 x N3 } -----------------------------------------
    for (int k = 0; k < e->N2; ++k) {
        stage_l2(e, st, 1, k);
        stage_l3_plain(e, st, 1, k);
    }
    stage_mid(e, st);
    for (int k = 0; k < e->N2; ++k) {
        stage_l2(e, st, 2, k);
        stage_l3_plain(e, st, 2, k);
    }
    stage_epilogue(e, st);
}

static int subcycle_rows(gm_engine* e, const float* rows, int n_rows) {
    if (n_rows != e->N23) return fail("subcycle needs exactly N2*N3 toolpath rows");
    GmStepDesc& h = e->host;
    h.nrows = n_rows;
    float dt1 = 0.0f;
    for (int i = 0; i < n_rows; ++i) {
        for (int c = 0; c < 7; ++c) h.rows[i][c] = rows[7 * i + c];
        dt1 += rows[7 * i + 5];
    }
    h.dt1 = dt1;
    for (int k = 0; k < e->N2; ++k) {
        float s2 = 0.0f;
        for (int j = 0; j < e->N3; ++j) s2 += rows[7 * (k * e->N3 + j) + 5];
        h.dt2[k] = s2;
    }
    return upload_desc(e, desc_bytes(n_rows));
}

extern "C" int gm_engine_subcycle(gm_engine* e, const float* rows, int n_rows) {
    if (e->stg_active) return fail("a staged subcycle is in progress");
    if (subcycle_rows(e, rows, n_rows)) return -1;
    if (!e->use_graph) {
        subcycle_body(e);
        HIPCHK(hipGetLastError());
        return 0;
    }
    if (!e->graph_ok) {
        if (e->gexec) {
            HIPCHK(hipGraphExecDestroy(e->gexec));
            e->gexec = nullptr;
        }
        hipGraph_t g;
        HIPCHK(hipStreamBeginCapture(e->s, hipStreamCaptureModeThreadLocal));
        subcycle_body(e);
        HIPCHK(hipStreamEndCapture(e->s, &g));
        HIPCHK(hipGraphInstantiate(&e->gexec, g, nullptr, nullptr, 0));
        HIPCHK(hipGraphDestroy(g));
        e->graph_ok = true;
    }
    HIPCHK(hipGraphLaunch(e->gexec, e->s));
    return 0;
}

// ---- the subcycled L1 step stage by stage ------------------------------------
// begin (rows, prologue) -> 2*N2 x { l2, l3 } with the middle and the
// epilogue run by the l3 stage that closes Part 1 and Part 2.  Kernels are
// launched directly (no graph), in the order gm_engine_subcycle records.

extern "C" int gm_engine_stage_begin(gm_engine* e, const float* rows, int n_rows) {
    if (e->stg_active) return fail("a staged subcycle is already in progress");
    if (subcycle_rows(e, rows, n_rows)) return -1;
    e->stg.T3last = nullptr;
    stage_prologue(e, e->stg);
    e->stg_active = 1;
    e->stg_part = 1;
    e->stg_k = 0;
    e->stg_phase = 0;
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_engine_stage_l2(gm_engine* e, int32_t* part, int32_t* k) {
    if (!e->stg_active || e->stg_phase != 0) return fail("gm_engine_stage_l2 out of order");
    stage_l2(e, e->stg, e->stg_part, e->stg_k);
    e->stg_phase = 1;
    if (part) *part = e->stg_part;
    if (k) *k = e->stg_k;
    HIPCHK(hipGetLastError());
    return 0;
}

// after an L3 stage: the next substep, the middle after Part 1, the
// epilogue after Part 2
static void stage_advance(gm_engine* e) {
    e->stg_phase = 0;
    if (++e->stg_k == e->N2) {
        e->stg_k = 0;
        if (e->stg_part == 1) {
            stage_mid(e, e->stg);
            e->stg_part = 2;
        } else {
            stage_epilogue(e, e->stg);
            e->stg_active = 0;
        }
    }
}

extern "C" int gm_engine_stage_l3(gm_engine* e, int external) {
    if (!e->stg_active || e->stg_phase != 1) return fail("gm_engine_stage_l3 out of order");
    if (external) stage_l3_external(e, e->stg, e->stg_part, e->stg_k);
    else stage_l3_plain(e, e->stg, e->stg_part, e->stg_k);
    stage_advance(e);
    HIPCHK(hipGetLastError());
    return 0;
}

// ---- many three-level engines with their L3 solves batched ---------------------
// The engines share one stream and run their subcycles stage by stage in
// lockstep; at each L2 substep the N3 L3 solves of all of them go out as N3
// launches of the batched step (one launch per solve for every engine).

// Pinned item slots of a gm_multi: the host may run this many substeps ahead
// of the GPU before it waits (a whole L1 step of N2 <= 16 substeps per part,
// so that the driver's host work for the next L1 step overlaps the GPU's
// substeps of this one).
constexpr int kMultiRing = 32;

struct gm_multi {
    int n = 0;
    hipStream_t s = nullptr;
    std::vector<gm_engine*> e;
    DevMat* d_mats = nullptr;
    BatchItem* d_items = nullptr;
    BatchItem* h_ring = nullptr;  // pinned, kMultiRing slots of the N3 x n items of a substep
    hipEvent_t ring_ev[kMultiRing] = {};
    int ring_slot = 0;
    int* d_bstart = nullptr;  // block starts of the 1 x 3 bricks (Part 2)
    int nblocks = 0;
    int tile1 = 1;               // brick shape of the Part-1 solves (kBatchTiles)
    int* d_bstart1 = nullptr;    // its block starts
    int nblocks1 = 0;
    bool rec = false;
    // the engines' other streams: the batched solves on s wait for their
    // L2 stages, their L3 closes wait for the batched solves
    std::vector<hipStream_t> xs;
    std::vector<hipEvent_t> xev;
    hipEvent_t mev = nullptr;
    // batched L2 stage and T' (multi_substep_batched): per substep one upload
    // of the L2-step, face and T' items and their block starts (slot layout
    // at off_*), from a pinned ring of kMultiRing slots
    bool batch_l2 = true;
    bool batch_restrict = true;  // the restrictions in the batched launch too
    int* d_bst_l2 = nullptr;
    int nblk_l2 = 0;
    char* d_sub2 = nullptr;
    char* h_ring2 = nullptr;
    hipEvent_t ring2_ev[kMultiRing] = {};
    int ring2_slot = 0;
    size_t sub2_bytes = 0, off_l2 = 0, off_face = 0, off_tp = 0, off_bf = 0, off_bt = 0;
    size_t off_ri = 0, off_br = 0;  // restriction items and their block starts
    // batched work on s not yet waited for by the other streams
    bool out_pending = false;
};

extern "C" int gm_multi_destroy(gm_multi* m) {
    if (!m) return 0;
    // wait for this object's own work only (the engines' streams may be gone):
    // the last batched launch and the item uploads
    if (m->mev) (void)hipEventSynchronize(m->mev);
    for (auto& ev : m->ring_ev)
        if (ev) (void)hipEventSynchronize(ev);
    for (hipEvent_t ev : m->xev)
        if (ev) (void)hipEventDestroy(ev);
    if (m->mev) (void)hipEventDestroy(m->mev);
    if (m->d_mats) (void)hipFree(m->d_mats);
    if (m->d_items) (void)hipFree(m->d_items);
    if (m->d_bstart) (void)hipFree(m->d_bstart);
    if (m->d_bstart1) (void)hipFree(m->d_bstart1);
    for (auto& ev : m->ring2_ev)
        if (ev) {
            (void)hipEventSynchronize(ev);
            (void)hipEventDestroy(ev);
        }
    if (m->h_ring2) (void)hipHostFree(m->h_ring2);
    if (m->d_sub2) (void)hipFree(m->d_sub2);
    if (m->d_bst_l2) (void)hipFree(m->d_bst_l2);
    if (m->h_ring) (void)hipHostFree(m->h_ring);
    for (auto& ev : m->ring_ev)
        if (ev) (void)hipEventDestroy(ev);
    delete m;
    return 0;
}

extern "C" int gm_multi_create(gm_engine* const* engines, int n, gm_multi** out) {
    if (n < 1 || !engines || !out) return fail("gm_multi_create: bad arguments");
    gm_multi* m = new gm_multi();
    auto bad = [&](int rc) {
        gm_multi_destroy(m);
        return rc;
    };
    m->n = n;
    m->e.assign(engines, engines + n);
    gm_engine* e0 = m->e[0];
    m->s = e0->s;
    m->rec = e0->dm.pnb > 0;
    std::vector<int> st(n + 1, 0);
    std::vector<DevMat> hm(n);
    for (int i = 0; i < n; ++i) {
        gm_engine* e = m->e[i];
        if (!e) return bad(fail("gm_multi_create: null engine"));
        if (std::find(m->e.begin(), m->e.begin() + i, e) != m->e.begin() + i)
            return bad(fail("gm_multi_create: an engine is listed twice"));
        if (e->s != m->s && std::find(m->xs.begin(), m->xs.end(), e->s) == m->xs.end())
            m->xs.push_back(e->s);
        if (e->N2 != e0->N2 || e->N3 != e0->N3)
            return bad(fail("gm_multi_create: the engines need the same N2 and N3"));
        if ((e->dm.pnb > 0) != m->rec)
            return bad(fail("gm_multi_create: all engines need the same kind of property tables"));
        const gm_grid& g = e->cfg.level[3];
        if ((int64_t)g.n[0] * g.n[1] * g.n[2] * 8 >= (int64_t)1 << 31)
            return bad(fail("gm_multi_create: L3 meshes must stay below 2^28 nodes"));
        hm[i] = e->dm;
        st[i + 1] = st[i] + batch_bricks(g, 1, 3);
    }
    m->nblocks = st[n];
#define MCHK(call)                                                                      \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return bad(fail(std::string(#call) + ": " + hipGetErrorString(e_)));        \
    } while (0)
    MCHK(hipMalloc(&m->d_mats, sizeof(DevMat) * n));
    MCHK(hipMalloc(&m->d_items, sizeof(BatchItem) * n * e0->N3));
    MCHK(hipMalloc(&m->d_bstart, sizeof(int) * (n + 1)));
    MCHK(hipMemcpy(m->d_mats, hm.data(), sizeof(DevMat) * n, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(m->d_bstart, st.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    MCHK(hipHostMalloc((void**)&m->h_ring, sizeof(BatchItem) * n * e0->N3 * kMultiRing,
                       hipHostMallocDefault));
    for (auto& ev : m->ring_ev) MCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    m->xev.assign(m->xs.size(), nullptr);
    for (auto& ev : m->xev) MCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    MCHK(hipEventCreateWithFlags(&m->mev, hipEventDisableTiming));
    {
        std::vector<int> b2(n + 1, 0);
        for (int i = 0; i < n; ++i) b2[i + 1] = b2[i] + batch_bricks(m->e[i]->cfg.level[2], 1, 2);
        m->nblk_l2 = b2[n];
        MCHK(hipMalloc(&m->d_bst_l2, sizeof(int) * (n + 1)));
        MCHK(hipMemcpy(m->d_bst_l2, b2.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
        size_t o = 0;
        m->off_l2 = o;
        o = align256(o + sizeof(BatchItem) * n);
        m->off_face = o;
        o = align256(o + sizeof(FaceItem) * n);
        m->off_tp = o;
        o = align256(o + sizeof(TpItem) * n);
        m->off_bf = o;
        o = align256(o + sizeof(int) * (n + 1));
        m->off_bt = o;
        o = align256(o + sizeof(int) * (n + 1));
        m->off_ri = o;
        o = align256(o + sizeof(RItem) * n);
        m->off_br = o;
        o = align256(o + sizeof(int) * (n + 1));
        m->sub2_bytes = o;
        MCHK(hipMalloc(&m->d_sub2, m->sub2_bytes));
        MCHK(hipHostMalloc((void**)&m->h_ring2, m->sub2_bytes * kMultiRing, hipHostMallocDefault));
        for (auto& ev : m->ring2_ev) MCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    }
#undef MCHK
    *out = m;
    return 0;
}

// the other streams' work so far before what m->s runs next
static int multi_join_in(gm_multi* m) {
    for (size_t i = 0; i < m->xs.size(); ++i) {
        HIPCHK(hipEventRecord(m->xev[i], m->xs[i]));
        HIPCHK(hipStreamWaitEvent(m->s, m->xev[i], 0));
    }
    return 0;
}

// m->s's work so far before what the other streams run next
static int multi_join_out(gm_multi* m) {
    HIPCHK(hipEventRecord(m->mev, m->s));
    for (hipStream_t x : m->xs) HIPCHK(hipStreamWaitEvent(x, m->mev, 0));
    return 0;
}

// the N3 batched L3 solves of substep (part, k) on m->s; w walks each
// engine's buffers
static int multi_l3_solves(gm_multi* m, int part, int k, std::vector<L3Walk>& w) {
    gm_engine* e0 = m->e[0];
    const int N3 = e0->N3, n = m->n;
    const float fN3 = (float)N3;
    // the items of all N3 solves, built in a pinned slot and uploaded at once
    const int slot = m->ring_slot;
    m->ring_slot = (slot + 1) % kMultiRing;
    HIPCHK(hipEventSynchronize(m->ring_ev[slot]));
    BatchItem* h = m->h_ring + (size_t)slot * n * N3;
    for (int j = 0; j < N3; ++j) {
        const int row = k * N3 + j;
        for (int i = 0; i < n; ++i) {
            gm_engine* e = m->e[i];
            const L3StepBufs b = l3_walk_step(e, e->stg, w[i], part, k, j);
            BatchItem& it = h[(size_t)j * n + i];
            it.a = l3_args_kc(e, b.Tin, b.S1in, b.S1out, row, (float)(j + 1) / fN3, b.Tout, b.S2in,
                              b.S2out);
            it.a.grid = e->cfg.level[3];  // node counts (a move changes the origin only)
            it.m = e->d_dm;
            for (int c = 0; c < 7; ++c) it.row[c] = e->host.rows[row][c];
            // the row's dt by value: the launch's desc is the first engine's
            it.a.dt_mode = GM_DT_VALUE;
            it.a.dt = it.row[5];
            it.w3 = e->w3[3];
        }
    }
    HIPCHK(hipMemcpyAsync(m->d_items, h, sizeof(BatchItem) * n * N3, hipMemcpyHostToDevice,
                          m->s));
    HIPCHK(hipEventRecord(m->ring_ev[slot], m->s));
    // brick shape of both parts' solves: the set one, else 1 x 3
    const bool t1 = m->d_bstart1 != nullptr;
    const int* bst = t1 ? m->d_bstart1 : m->d_bstart;
    const int nbk = t1 ? m->nblocks1 : m->nblocks, tile = t1 ? m->tile1 : 1;
    for (int j = 0; j < N3; ++j) {
        const BatchItem* it = m->d_items + (size_t)j * n;
        if (part == 1)
            launch_batch_step(it, m->d_mats, bst, n, nbk, tile, m->rec, e0->d_desc, m->s);
        else
            launch_batch_step_acc(it, m->d_mats, bst, n, nbk, tile, m->rec, e0->d_desc, m->s);
    }
    return 0;
}

extern "C" int gm_multi_l3(gm_multi* m) {
    if (!m) return fail("gm_multi_l3: null");
    gm_engine* e0 = m->e[0];
    const int part = e0->stg_part, k = e0->stg_k;
    for (gm_engine* e : m->e)
        if (!e->stg_active || e->stg_phase != 1 || e->stg_part != part || e->stg_k != k)
            return fail("gm_multi_l3: every engine must be between the l2 and l3 stages of "
                        "the same substep");
    std::vector<L3Walk> w(m->n);
    for (int i = 0; i < m->n; ++i) w[i] = l3_walk_begin(m->e[i]->stg);
    if (int rc = multi_join_in(m)) return rc;
    if (int rc = multi_l3_solves(m, part, k, w)) return rc;
    if (int rc = multi_join_out(m)) return rc;
    for (int i = 0; i < m->n; ++i) {
        gm_engine* e = m->e[i];
        stage_l3_close(e, e->stg, part, k, w[i].S13in, w[i].S23in, e->stg.T3out);
        stage_advance(e);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

// One whole substep of every engine with the L2 stage and T' batched too
// (engines on the node-aligned path): on m->s one launch each for the
// restrictions of the L3 T' terms, the L2 solves, the face values, the N3 L3
// solves and T' + injection.  An engine whose restriction the batched launch
// does not run restricts on its own stream, and the streams are joined around
// it; otherwise m->s waits for the other streams only at the first substep of
// each part (after the prologue or the middle of the L1 step) and they wait
// for it after the last (before the middle or the epilogue).
static int multi_substep_batched(gm_multi* m) {
    gm_engine* e0 = m->e[0];
    const int part = e0->stg_part, k = e0->stg_k, n = m->n;
    const int slot = m->ring2_slot;
    m->ring2_slot = (slot + 1) % kMultiRing;
    HIPCHK(hipEventSynchronize(m->ring2_ev[slot]));
    char* h = m->h_ring2 + (size_t)slot * m->sub2_bytes;
    BatchItem* l2 = (BatchItem*)(h + m->off_l2);
    FaceItem* fi = (FaceItem*)(h + m->off_face);
    TpItem* tp = (TpItem*)(h + m->off_tp);
    int* bf = (int*)(h + m->off_bf);
    int* bt = (int*)(h + m->off_bt);
    RItem* ri = (RItem*)(h + m->off_ri);
    int* br = (int*)(h + m->off_br);
    bf[0] = bt[0] = br[0] = 0;
    int nri = 0;
    bool own = false;  // an engine restricted on its own stream
    for (int i = 0; i < n; ++i) {
        gm_engine* e = m->e[i];
        SubcycleState& st = e->stg;
        const RArgs ra = stage_l2_rargs(e, st, part, k);
        if (m->batch_restrict && restrict_item(e, 2, ra, &ri[nri])) {
            br[nri + 1] = br[nri] + restrict_batch_blocks(ri[nri].p, ri[nri].r);
            ++nri;
        } else {
            // its inputs come from the batched work on m->s
            if (m->out_pending) {
                if (int rc = multi_join_out(m)) return rc;
                m->out_pending = false;
            }
            eng_restrict(e, 2, ra);
            own = true;
        }
        l2[i].a = stage_l2_args(e, st, k);
        l2[i].m = e->d_dm;
        memset(l2[i].row, 0, sizeof(l2[i].row));
        l2[i].w3 = e->w3[2];
        fi[i] = FaceItem{e->d_desc, st.T2new, st.T2cur, e->fa3, e->fb3};
        bf[i + 1] = bf[i] + face_bc_blocks(e->cfg.level[3]);
        stage_l2_post(e, st, part, k);
        int nb = 0;
        tp[i] = stage_tprime_item(e, st, k, st.T3out, &nb);
        bt[i + 1] = bt[i] + nb;
        e->stg_phase = 1;
    }
    if (k == 0 || own)
        if (int rc = multi_join_in(m)) return rc;
    HIPCHK(hipMemcpyAsync(m->d_sub2, h, m->sub2_bytes, hipMemcpyHostToDevice, m->s));
    HIPCHK(hipEventRecord(m->ring2_ev[slot], m->s));
    char* d = m->d_sub2;
    // all items of one part have the same terms: flux (Part 1), flux + mass
    launch_restrict_batch((const RItem*)(d + m->off_ri), (const int*)(d + m->off_br), nri,
                          br[nri], true, part == 2, m->s);
    launch_batch_step_l2((const BatchItem*)(d + m->off_l2), m->d_mats, m->d_bst_l2, n,
                         m->nblk_l2, m->rec, e0->d_desc, m->s);
    launch_face_bc_batch((const FaceItem*)(d + m->off_face), (const int*)(d + m->off_bf), n,
                         bf[n], m->s);
    std::vector<L3Walk> w(n);
    for (int i = 0; i < n; ++i) w[i] = l3_walk_begin(m->e[i]->stg);
    if (int rc = multi_l3_solves(m, part, k, w)) return rc;
    launch_tprime_batch((const TpItem*)(d + m->off_tp), (const int*)(d + m->off_bt), n, bt[n],
                        m->s);
    m->out_pending = true;
    if (k == e0->N2 - 1 || own) {
        if (int rc = multi_join_out(m)) return rc;
        m->out_pending = false;
    }
    for (int i = 0; i < n; ++i) {
        gm_engine* e = m->e[i];
        stage_l3_close_state(e, e->stg, part, k, w[i].S13in, w[i].S23in, e->stg.T3out);
        stage_advance(e);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_multi_set_batch_l2(gm_multi* m, int on) {
    if (!m) return fail("gm_multi_set_batch_l2: null");
    m->batch_l2 = on != 0;
    return 0;
}

extern "C" int gm_multi_set_batch_restrict(gm_multi* m, int on) {
    if (!m) return fail("gm_multi_set_batch_restrict: null");
    m->batch_restrict = on != 0;
    return 0;
}

extern "C" int gm_multi_set_tile(gm_multi* m, int yb, int zc) {
    if (!m) return fail("gm_multi_set_tile: null");
    int z = -1;
    for (int t = 0; t < GM_BATCH_TILES; ++t)
        if (kBatchTiles[t][0] == yb && kBatchTiles[t][1] == zc) z = t;
    if (z < 0)
        return fail("gm_multi_set_tile: (rows, planes) must be one of (1,2) (1,3) (1,4) (2,2) (2,3)");
    std::vector<int> st(m->n + 1, 0);
    for (int i = 0; i < m->n; ++i)
        st[i + 1] = st[i] + batch_bricks(m->e[i]->cfg.level[3], yb, zc);
    // the previous table may still be read by launches in flight
    HIPCHK(hipStreamSynchronize(m->s));
    if (!m->d_bstart1) HIPCHK(hipMalloc(&m->d_bstart1, sizeof(int) * (m->n + 1)));
    HIPCHK(hipMemcpy(m->d_bstart1, st.data(), sizeof(int) * (m->n + 1), hipMemcpyHostToDevice));
    m->tile1 = z;
    m->nblocks1 = st[m->n];
    return 0;
}

extern "C" int gm_multi_subcycle(gm_multi* m) {
    if (!m) return fail("gm_multi_subcycle: null");
    for (gm_engine* e : m->e)
        if (!e->stg_active || e->stg_phase != 0 || e->stg_part != 1 || e->stg_k != 0)
            return fail("gm_multi_subcycle: every engine must have just begun a staged subcycle "
                        "(gm_engine_stage_begin)");
    const int n2 = m->e[0]->N2;
    bool batched = m->batch_l2 && m->d_sub2;
    for (gm_engine* e : m->e) batched = batched && e->fast_restrict && e->use_kc;
    for (int sub = 0; sub < 2 * n2; ++sub) {
        if (batched) {
            if (int rc = multi_substep_batched(m)) return rc;
            continue;
        }
        for (gm_engine* e : m->e) {
            stage_l2(e, e->stg, e->stg_part, e->stg_k);
            e->stg_phase = 1;
        }
        if (int rc = gm_multi_l3(m)) return rc;
    }
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_engine_stage_field(gm_engine* e, int which, void** ptr, int64_t* count) {
    const SubcycleState& st = e->stg;
    if (which == GM_STAGE_L3_LAST) {
        if (!st.T3last) return fail("gm_engine_stage_field: no L3 stage has run yet");
        *ptr = (void*)st.T3last;
        *count = e->nn[3];
        return 0;
    }
    if (!e->stg_active || e->stg_phase != 1)
        return fail("gm_engine_stage_field: only between the l2 and l3 stages");
    void* p = nullptr;
    switch (which) {
        case GM_STAGE_L3_IN: p = (void*)st.T3cur; break;
        case GM_STAGE_L3_OUT: p = (void*)st.T3ext; break;
        case GM_STAGE_FACE_NEW: p = (void*)e->fa3; break;
        case GM_STAGE_FACE_OLD: p = (void*)e->fb3; break;
        case GM_STAGE_S1_IN: p = (void*)st.S13cur; break;
        default: return fail("gm_engine_stage_field: no such field");
    }
    *ptr = p;
    *count = e->nn[3];
    return 0;
}

// ---- stepGOMELT (cF:2304-2397) ----------------------------------------------

// _resetmask of cF:2394-2397: (1 - 2 S2_before) * S2_after == 1.
__global__ void k_reset_mask(const uint8_t* __restrict__ pre, const uint8_t* __restrict__ post,
                             uint8_t* __restrict__ reset, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) reset[t] = (uint8_t)((pre[t] == 0) && (post[t] != 0));
}

static void solutions(gm_engine* e, const float* V1, const float* V2, float* T1o, float* T2o,
                      float* T3o) {
    // computeSolutions (cF:2135-2204): L1, then L2 from the new L1, then L3
    // from the new L2; no clamp inside (the caller clamps afterwards).
    GmStepDesc* d = e->d_desc;
    run_l1(e, V1, T1o, GM_DT_ROW, 0.0f, 0);
    launch_face_bc(d, 2, 1, e->cfg.level[2], T1o, T1o, e->fa2, e->fb2, e->s);
    run_l2(e, e->T2, e->S12, nullptr, V2, 1.0f, T2o, GM_DT_ROW, 0, 0.0f, 0);
    launch_face_bc(d, 3, 2, e->cfg.level[3], T2o, T2o, e->fa3, e->fb3, e->s);
    run_l3(e, e->T3, e->S13, nullptr, 0, 1.0f, T3o, 0, nullptr, nullptr);
}

extern "C" int gm_engine_step(gm_engine* e, const float* row) {
    GmStepDesc& h = e->host;
    h.nrows = 1;
    for (int c = 0; c < 7; ++c) h.rows[0][c] = row[c];
    h.dt1 = row[5];
    if (upload_desc(e, desc_bytes(1))) return -1;
    GmStepDesc* d = e->d_desc;
    hipStream_t s = e->s;
    const int64_t n1 = e->nn[1], n2 = e->nn[2], n3 = e->nn[3];
    const float Tamb = e->mat.T_amb;
    launch_copy_u8(e->preS2, e->S23, n3, s);
    // updateStateProperties (cF:2513-2564): persist L3 S1/S2 and L2 S1
    extern void gm_props_u8(const DevMat& m, const float* T, uint8_t* S1, uint8_t* S2, int64_t n,
                            int64_t nsub, hipStream_t s);
    gm_props_u8(e->dm, e->T3, e->S13, e->S23, n3, h.sub[3], s);
    gm_props_u8(e->dm, e->T2, e->S12, nullptr, n2, h.sub[2], s);
    l1_s1_update(e);
    // sources projected to L1 and L2 (computeSources cF:928-988)
    launch_source_tables(e->pair[1].p, d, e->dm, 0, 1, 0, nullptr, e->G1, e->gmax[1], s);
    launch_source_tables(e->pair[2].p, d, e->dm, 0, 1, 0, nullptr, e->G2, e->gmax[2], s);
    // predictor T' flux terms (computeCoarseTprimeTerm_jax cF:1477-1565)
    eng_restrict(e, 1,
                    rargs(e, 3, e->T3, e->S13, e->Tp3, nullptr, nullptr, GM_DT_VALUE, 0, 1.0f));
    eng_restrict(e, 0,
                    rargs(e, 2, e->T2, e->S12, e->Tp2, nullptr, nullptr, GM_DT_VALUE, 0, 1.0f));
    eng_restrict_nodes(e, 0, 1, nullptr, e->V1,
                          1, e->G1, 1, e->gmax[1], e->w3[3]);
    eng_restrict(e, 2,
                    rargs(e, 3, e->T3, e->S13, e->Tp3, nullptr, nullptr, GM_DT_VALUE, 0, 1.0f));
    eng_restrict_nodes(e, 2, -1, nullptr, e->V2,
                          2, e->G2, 1, e->gmax[2], e->w3[3]);
    // predictor
    float* T1o = e->T1p;
    float* T2o = e->T2b[0];
    float* T3o = e->T3b[0];
    solutions(e, e->V1, e->V2, T1o, T2o, T3o);
    launch_clamp(T1o, n1, Tamb, s);
    launch_clamp(T2o, n2, Tamb, s);
    launch_clamp(T3o, n3, Tamb, s);
    new_tprime(e, 3, T3o, T2o, e->Tp3x);
    new_tprime(e, 2, T2o, T1o, e->Tp2x);
    // mass terms (computeCoarseTprimeMassTerm_jax cF:1396-1474)
    eng_restrict(e, 1,
                    rargs(e, 3, e->T3, e->S13, nullptr, e->Tp3x, e->Tp3, GM_DT_ROW, 0, 0.0f));
    eng_restrict(e, 0,
                    rargs(e, 2, e->T2, e->S12, nullptr, e->Tp2x, e->Tp2, GM_DT_ROW, 0, 0.0f));
    eng_restrict_nodes(e, 0, 1, e->V1, e->V1m,
                          -1, nullptr, 0, 0, 0.0f);
    eng_restrict(e, 2,
                    rargs(e, 3, e->T3, e->S13, nullptr, e->Tp3x, e->Tp3, GM_DT_ROW, 0, 0.0f));
    float* V2m = e->Tp2x;  // reuse: the L2 T' of the predictor is no longer needed
    eng_restrict_nodes(e, 2, -1, e->V2, V2m, -1, nullptr,
                          0, 0, 0.0f);
    // corrector
    T1o = e->T1c;
    T2o = e->T2b[1];
    T3o = e->T3b[1];
    solutions(e, e->V1m, V2m, T1o, T2o, T3o);
    launch_clamp(T1o, n1, Tamb, s);
    launch_clamp(T2o, n2, Tamb, s);
    launch_clamp(T3o, n3, Tamb, s);
    // final T' (getBothNewTprimes cF:2385-2387)
    launch_copy_f32(e->T3, T3o, n3, s);
    new_tprime(e, 3, e->T3, T2o, e->Tp3);
    new_tprime(e, 2, T2o, T1o, e->Tp2);
    launch_copy_f32(e->T2, T2o, n2, s);
    launch_copy_f32(e->T1, T1o, n1, s);
    // Level0 state and the driver's melt-time bookkeeping (go_melt.py:339-357)
    launch_l0_scatter(d, e->S13, e->S23, e->L0S1, e->s2snap, e->s2box, n3, s);
    // the reset mask stepGOMELT returns (cF:2394-2397): S2 went 0 -> 1
    hipLaunchKernelGGL(k_reset_mask, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, s,
                       e->preS2, e->S23, e->reset3, n3);
    if (!e->acc_external)
        launch_melt_single(d, e->T3, e->preS2, e->S23, row[5], e->mat.T_liquidus, e->acc,
                           e->macc, n3, s);
    HIPCHK(hipGetLastError());
    return 0;
}

// computeStateProperties on uint8 state, persisting S1 (and S2).
__global__ void k_props_u8(DevMat m, const float* __restrict__ T, uint8_t* S1, uint8_t* S2,
                           int64_t n, int64_t nsub) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float v = T[t];
    const bool s2 = v >= m.T_liq;
    S1[t] = (uint8_t)((S1[t] != 0) || s2 || t < nsub);
    if (S2) S2[t] = (uint8_t)s2;
}

void gm_props_u8(const DevMat& m, const float* T, uint8_t* S1, uint8_t* S2, int64_t n,
                 int64_t nsub, hipStream_t s) {
    hipLaunchKernelGGL(k_props_u8, dim3((n + 255) / 256), dim3(256), 0, s, m, T, S1, S2, n, nsub);
}

// ---- stepGOMELTDwellTime (cF:2617-2664) -------------------------------------

extern "C" int gm_engine_step_dwell(gm_engine* e, float dt) {
    run_l1(e, nullptr, e->T1p, GM_DT_VALUE, dt, 0);
    launch_copy_f32(e->T1, e->T1p, e->nn[1], e->s);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_engine_clear_tprime(gm_engine* e) {
    HIPCHK(hipMemsetAsync(e->Tp2, 0, e->nn[2] * 4, e->s));
    HIPCHK(hipMemsetAsync(e->Tp3, 0, e->nn[3] * 4, e->s));
    return 0;
}

extern "C" int gm_engine_level0_substrate(gm_engine* e, int64_t n) {
    launch_fill_u8(e->L0S1, 1, std::min(n, e->nn[0]), e->s);
    HIPCHK(hipGetLastError());
    return 0;
}

// ---- layer change of the single-step driver path (go_melt.py:198-290) -------

extern "C" int gm_engine_layer_change(gm_engine* e, int kidx, float new_l1_z0, float laser_z,
                                      int shift_level0) {
    (void)laser_z;
    GmStepDesc& h = e->host;
    hipStream_t s = e->s;
    const int64_t n1 = e->nn[1];
    // T0 re-interpolated onto the shifted L1 z axis, clamped at T_amb
    gm_grid newg = h.g[1];
    newg.x0[2] = new_l1_z0;
    launch_interp_grid(nullptr, -1, &h.g[1], e->T1, nullptr, -1, &newg, n1, e->T1p, s);
    launch_clamp(e->T1p, n1, e->mat.T_amb, s);
    launch_copy_f32(e->T1, e->T1p, n1, s);
    // S1 storage: store the current S1 in row kidx-1 and load row kidx
    // (indices outside the storage clip to its ends, as JAX indexing does)
    const int st = e->nst;
    int wrow = kidx - 1;
    if (wrow < 0) wrow += st;
    wrow = std::min(std::max(wrow, 0), st - 1);
    const int rrow = std::min(std::max(kidx, 0), st - 1);
    launch_copy_f32(e->S1st + (size_t)wrow * n1, e->S11, n1, s);
    launch_copy_f32(e->S11, e->S1st + (size_t)rrow * n1, n1, s);
    h.g[1].x0[2] = new_l1_z0;
    h.gprev[1] = h.g[1];
    if (shift_level0) {
        const gm_grid& g0 = e->cfg.level[0];
        const int64_t nxy = (int64_t)g0.n[0] * g0.n[1];
        const int64_t nA = nxy * e->cfg.l0_layer_delta;
        const int64_t nB = nxy * (g0.n[2] - e->cfg.l0_layer_delta);
        // accum = max(accum, max_accum), then shift down and clear the maxima
        extern void gm_vmax(float* a, const float* b, int64_t n, hipStream_t s);
        gm_vmax(e->acc, e->macc, e->nn[0], s);
        if (nA > 0) launch_shift_l0(e->L0S1, e->acc, nA, nB, e->nn[0], s);
        HIPCHK(hipMemsetAsync(e->macc, 0, e->nn[0] * 4, s));
    }
    e->graph_ok = false;
    if (update_alignment(e)) return -1;
    if (upload_desc(e, desc_bytes(0))) return -1;
    HIPCHK(hipGetLastError());
    return 0;
}

__global__ void k_vmax(float* a, const float* b, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n && b[t] > a[t]) a[t] = b[t];
}

void gm_vmax(float* a, const float* b, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_vmax, dim3((n + 255) / 256), dim3(256), 0, s, a, b, n);
}

// Times `iters` back-to-back launches of one kernel of the L1 step with the
// engine's current state (the Part-1 arguments of the first substep), recorded
// as a hipGraph and bracketed by HIP events on the engine stream.
//   which = 3: L3 level step, 2: L2 level step, 1: L1 level step,
//           4: L3 -> L2 restriction (elements + nodes)
extern "C" int gm_engine_time_kernel(gm_engine* e, int which, int iters, float* ms_per_launch) {
    if (iters < 1) return fail("iters must be >= 1");
    if (which == 0) {
        // replays of the recorded subcycle graph (the last L1 step's rows)
        if (!e->graph_ok || !e->gexec) return fail("no recorded subcycle graph to time");
        hipEvent_t a, b;
        HIPCHK(hipEventCreate(&a));
        HIPCHK(hipEventCreate(&b));
        HIPCHK(hipGraphLaunch(e->gexec, e->s));  // warm
        HIPCHK(hipEventRecord(a, e->s));
        for (int i = 0; i < iters; ++i) HIPCHK(hipGraphLaunch(e->gexec, e->s));
        HIPCHK(hipEventRecord(b, e->s));
        HIPCHK(hipEventSynchronize(b));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, a, b));
        *ms_per_launch = ms / (float)iters;
        HIPCHK(hipEventDestroy(a));
        HIPCHK(hipEventDestroy(b));
        return 0;
    }
    hipGraph_t g;
    hipGraphExec_t ge;
    HIPCHK(hipStreamSynchronize(e->s));
    HIPCHK(hipStreamBeginCapture(e->s, hipStreamCaptureModeThreadLocal));
    e->use_kc = true;
    for (int i = 0; i < iters; ++i) {
        const int b = i & 1;
        if (which == 3)
            run_l3_kc(e, b ? e->T3b[0] : e->T3, e->S13, e->S13b[b], 0, 1.0f, e->T3b[b ? 1 : 0],
                      nullptr, nullptr);
        else if (which == 2)
            run_l2(e, b ? e->T2b[0] : e->T2, e->S12, e->S12b[b], e->V2, 1.0f, e->T2b[b ? 1 : 0],
                   GM_DT_L2, 0, 0.0f, 1);
        else if (which == 1)
            run_l1(e, e->V1, b ? e->T1c : e->T1p, GM_DT_L1, 0.0f, 1);
        else if (which == 4) {
            eng_restrict(e, 2,
                            rargs(e, 3, e->T3, e->S13, e->Tp3, e->histTp[0], e->Tp3, GM_DT_L2, 0,
                                  0.0f));
            eng_restrict_nodes(e, 2, -1, nullptr, e->V2,
                                  -1, nullptr, 0, 0, 0.0f);
        }
    }
    e->use_kc = false;
    HIPCHK(hipStreamEndCapture(e->s, &g));
    HIPCHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipGraphLaunch(ge, e->s));  // warm
    HIPCHK(hipEventRecord(a, e->s));
    HIPCHK(hipGraphLaunch(ge, e->s));
    HIPCHK(hipEventRecord(b, e->s));
    HIPCHK(hipEventSynchronize(b));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, a, b));
    *ms_per_launch = ms / (float)iters;
    HIPCHK(hipEventDestroy(a));
    HIPCHK(hipEventDestroy(b));
    HIPCHK(hipGraphExecDestroy(ge));
    HIPCHK(hipGraphDestroy(g));
    return 0;
}

extern "C" int gm_engine_minmax(gm_engine* e, float* out) {
    launch_minmax(e->T1, e->nn[1], e->mm, e->s);
    launch_minmax(e->T2, e->nn[2], e->mm + 2, e->s);
    launch_minmax(e->T3, e->nn[3], e->mm + 4, e->s);
    HIPCHK(hipMemcpyAsync(out, e->mm, 6 * 4, hipMemcpyDeviceToHost, e->s));
    HIPCHK(hipStreamSynchronize(e->s));
    return 0;
}

// ---------------------------------------------------------------------------
// Batched instances: many independent L3-type windows of their own mesh size
// and laser, advanced by one launch per step (k_step_dpp_batch)
// ---------------------------------------------------------------------------

struct gm_batch {
    int n = 0;
    hipStream_t s = nullptr;
    bool rec = false;
    int tile = 1;                     // brick shape of the batched launch (kBatchTiles)
    int cur = 0;                      // T[cur] holds the current field
    std::vector<gm_grid> g;
    DevMat* d_mats = nullptr;         // one per instance (laser constants differ)
    std::vector<DevMat*> dm;          // d_mats + i
    std::vector<float*> T[2], kc[2];  // fields and their carried k | cp
    std::vector<uint8_t*> S1, S1t;    // S1 (updated in place) and a tuning scratch
    std::vector<float*> face, src;    // Dirichlet face values, source tables
    std::vector<BatchItem> items;     // host copy of the step arguments
    BatchItem* d_items = nullptr;
    BatchItem* h_ring = nullptr;      // pinned upload slots
    hipEvent_t ring_ev[4] = {};
    int ring_slot = 0;
    int* d_bstart[GM_BATCH_TILES] = {};  // block starts per brick shape
    int nblocks[GM_BATCH_TILES] = {};
    int maxaxes = 0;
    GmStepDesc* d_desc = nullptr;     // zeroed; the batched step reads no row from it
    std::vector<void*> allocs;
};


extern "C" int gm_batch_destroy(gm_batch* b) {
    if (!b) return 0;
    if (b->s) (void)hipStreamSynchronize(b->s);
    for (void* p : b->allocs) (void)hipFree(p);
    if (b->h_ring) (void)hipHostFree(b->h_ring);
    for (auto& ev : b->ring_ev)
        if (ev) (void)hipEventDestroy(ev);
    delete b;
    return 0;
}

template <class P>
static hipError_t balloc(gm_batch* b, P** p, size_t bytes) {
    hipError_t rc = hipMalloc((void**)p, bytes ? bytes : 1);
    if (rc == hipSuccess) b->allocs.push_back((void*)*p);
    return rc;
}

extern "C" int gm_batch_create(int n, const gm_grid* grids, const gm_material* mats,
                               void* stream, gm_batch** out) {
    if (n < 1 || !grids || !mats || !out) return fail("gm_batch_create: bad arguments");
    gm_batch* b = new gm_batch();
    auto bad = [&](int rc) {
        gm_batch_destroy(b);
        return rc;
    };
#define BCHK(call)                                                                      \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return bad(fail(std::string(#call) + ": " + hipGetErrorString(e_)));        \
    } while (0)
    b->n = n;
    b->s = (hipStream_t)stream;
    b->g.assign(grids, grids + n);
    const DevMat hm0 = make_devmat(mats[0]);
    b->rec = hm0.pnb > 0;
    for (int i = 0; i < n; ++i) {
        const gm_grid& g = grids[i];
        if (g.n[0] < 2 || g.n[1] < 2 || g.n[2] < 2 ||
            (int64_t)g.n[0] * g.n[1] * g.n[2] * 8 >= (int64_t)1 << 31)
            return bad(fail("gm_batch_create: instance meshes need >= 2 nodes per axis and < 2^28 nodes"));
        if ((make_devmat(mats[i]).pnb > 0) != b->rec)
            return bad(fail("gm_batch_create: all instances need the same kind of property tables"));
        b->maxaxes = std::max(b->maxaxes, g.n[0] + g.n[1] + g.n[2]);
    }
    b->dm.resize(n);
    BCHK(balloc(b, &b->d_mats, sizeof(DevMat) * n));
    for (int k = 0; k < 2; ++k) {
        b->T[k].resize(n);
        b->kc[k].resize(n);
    }
    b->S1.resize(n);
    b->S1t.resize(n);
    b->face.resize(n);
    b->src.resize(n);
    b->items.resize(n);
    for (int i = 0; i < n; ++i) {
        const gm_grid& g = grids[i];
        const int64_t nn = (int64_t)g.n[0] * g.n[1] * g.n[2];
        const DevMat hm = make_devmat(mats[i]);
        b->dm[i] = b->d_mats + i;
        BCHK(hipMemcpyAsync(b->dm[i], &hm, sizeof(DevMat), hipMemcpyHostToDevice, b->s));
        for (int k = 0; k < 2; ++k) {
            BCHK(balloc(b, &b->T[k][i], nn * 4));
            BCHK(balloc(b, &b->kc[k][i], 2 * nn * 4));
        }
        BCHK(balloc(b, &b->S1[i], nn));
        BCHK(balloc(b, &b->S1t[i], nn));
        BCHK(balloc(b, &b->face[i], nn * 4));
        BCHK(balloc(b, &b->src[i], (size_t)(g.n[0] + g.n[1] + g.n[2]) * 4));
        launch_fill_f32(b->T[0][i], mats[i].T_amb, nn, b->s);
        launch_fill_f32(b->face[i], mats[i].T_amb, nn, b->s);
        launch_fill_u8(b->S1[i], 0, nn, b->s);
        launch_props_kc(b->dm[i], b->T[0][i], b->S1[i], nn, 0, b->kc[0][i], b->s);
        LSArgs a = ls_default();
        a.use_desc = 0;
        a.level = 3;
        a.grid = g;
        a.n_sub = 0;
        a.nn_active = nn;
        a.ne_active = (int64_t)(g.n[0] - 1) * (g.n[1] - 1) * (g.n[2] - 1);
        a.dt_mode = GM_DT_VALUE;
        a.src_mode = GM_SRC_L3;
        a.src_row = 0;
        a.srcT = b->src[i];
        a.conv = 1;
        a.bcA = a.bcB = b->face[i];
        a.alpha = 1.0f;
        a.clamp = 1;
        // carried k, cp are in the raw units of the material tables: k / 1000
        // and rho fold into the coefficients, as for the engine's levels
        ls_coeffs(a, g);
        for (int q = 0; q < 8; ++q) a.kc[q] = (float)((double)a.kc[q] / 1000.0);
        a.mrow = (float)((double)a.mrow * mats[i].rho);
        a.prec_ok = b->rec;
        b->items[i].a = a;
        b->items[i].m = b->dm[i];
        b->items[i].w3 = a.w3;
        memset(b->items[i].row, 0, sizeof(b->items[i].row));
    }
    BCHK(balloc(b, &b->d_items, sizeof(BatchItem) * n));
    BCHK(balloc(b, &b->d_desc, sizeof(GmStepDesc)));
    BCHK(hipMemsetAsync(b->d_desc, 0, sizeof(GmStepDesc), b->s));
    for (int z = 0; z < GM_BATCH_TILES; ++z) {
        std::vector<int> st(n + 1, 0);
        for (int i = 0; i < n; ++i)
            st[i + 1] = st[i] + batch_bricks(grids[i], kBatchTiles[z][0], kBatchTiles[z][1]);
        b->nblocks[z] = st[n];
        BCHK(balloc(b, &b->d_bstart[z], sizeof(int) * (n + 1)));
        BCHK(hipMemcpyAsync(b->d_bstart[z], st.data(), sizeof(int) * (n + 1),
                              hipMemcpyHostToDevice, b->s));
    }
    if (hipHostMalloc((void**)&b->h_ring, sizeof(BatchItem) * n * 4, hipHostMallocDefault) !=
        hipSuccess) {
        b->h_ring = nullptr;
        return bad(fail("gm_batch_create: hipHostMalloc failed"));
    }
    for (auto& ev : b->ring_ev) BCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    BCHK(hipStreamSynchronize(b->s));
    *out = b;
    return 0;
}
#undef BCHK

// Step arguments of every instance for one step: row i of rows (x, y, z,
// Ljump, Ldwell, dt, P), input T[cur], output T[cur ^ 1]; S1 in place or
// into the tuning scratch.  Uploaded through a pinned slot.
static int batch_upload(gm_batch* b, const float* rows, bool scratch_s1) {
    for (int i = 0; i < b->n; ++i) {
        BatchItem& it = b->items[i];
        if (rows) memcpy(it.row, rows + 7 * i, sizeof(it.row));
        LSArgs& a = it.a;
        a.dt = it.row[5];
        a.T = b->T[b->cur][i];
        a.S1 = b->S1[i];
        a.S1out = scratch_s1 ? b->S1t[i] : b->S1[i];
        a.Tout = b->T[b->cur ^ 1][i];
        a.kcin = b->kc[b->cur][i];
        a.kcout = b->kc[b->cur ^ 1][i];
    }
    const int slot = b->ring_slot;
    b->ring_slot = (slot + 1) & 3;
    HIPCHK(hipEventSynchronize(b->ring_ev[slot]));
    BatchItem* h = b->h_ring + (size_t)slot * b->n;
    memcpy(h, b->items.data(), sizeof(BatchItem) * b->n);
    HIPCHK(hipMemcpyAsync(b->d_items, h, sizeof(BatchItem) * b->n, hipMemcpyHostToDevice, b->s));
    HIPCHK(hipEventRecord(b->ring_ev[slot], b->s));
    return 0;
}

static void batch_launch(gm_batch* b, int batched) {
    launch_batch_src(b->d_items, b->n, b->maxaxes, b->s);
    if (batched) {
        const int z = b->tile;
        launch_batch_step(b->d_items, b->d_mats, b->d_bstart[z], b->n, b->nblocks[z], z, b->rec,
                          b->d_desc, b->s);
    } else {
        for (int i = 0; i < b->n; ++i)
            launch_step_tile(b->items[i].a, b->d_desc, b->g[i], b->dm[i], b->tile, b->rec, b->s);
    }
}

extern "C" int gm_batch_step(gm_batch* b, const float* rows_host, int batched) {
    if (!b || !rows_host) return fail("gm_batch_step: null argument");
    if (int rc = batch_upload(b, rows_host, false)) return rc;
    batch_launch(b, batched);
    HIPCHK(hipGetLastError());
    b->cur ^= 1;
    return 0;
}

extern "C" int gm_batch_set_tile(gm_batch* b, int yb, int zc) {
    if (!b) return fail("gm_batch_set_tile: null batch");
    for (int t = 0; t < GM_BATCH_TILES; ++t)
        if (kBatchTiles[t][0] == yb && kBatchTiles[t][1] == zc) {
            b->tile = t;
            return 0;
        }
    return fail("gm_batch_set_tile: (rows, planes) must be one of (1,2) (1,3) (1,4) (2,2) (2,3)");
}

// Times `iters` steps from the current state (outputs into the next buffers
// and the S1 scratch, so the state does not advance), batched or not.
extern "C" int gm_batch_time(gm_batch* b, int batched, int iters, float* ms_per_step) {
    if (!b || iters < 1 || !ms_per_step) return fail("gm_batch_time: bad arguments");
    if (int rc = batch_upload(b, nullptr, true)) return rc;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    batch_launch(b, batched);  // warm-up
    HIPCHK(hipEventRecord(e0, b->s));
    for (int k = 0; k < iters; ++k) batch_launch(b, batched);
    HIPCHK(hipEventRecord(e1, b->s));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCHK(hipGetLastError());
    *ms_per_step = ms / iters;
    return 0;
}

// Autotuner: the planes per brick (2, 3 or 4) of the batched launch with the
// shortest measured step on this batch; ms_host (nullable) gets the three
// times.  The state does not advance.
extern "C" int gm_batch_autotune(gm_batch* b, int iters, float* ms_host, int* yb_out,
                                 int* zc_out) {
    if (!b || iters < 1) return fail("gm_batch_autotune: bad arguments");
    float best = 0.0f;
    int bt = 1;
    for (int t = 0; t < GM_BATCH_TILES; ++t) {
        b->tile = t;
        float ms = 0.0f;
        if (int rc = gm_batch_time(b, 1, iters, &ms)) return rc;
        if (ms_host) ms_host[t] = ms;
        if (t == 0 || ms < best) {
            best = ms;
            bt = t;
        }
    }
    b->tile = bt;
    if (yb_out) *yb_out = kBatchTiles[bt][0];
    if (zc_out) *zc_out = kBatchTiles[bt][1];
    return 0;
}

// Device views of instance i: field 0 = T (FP32), 1 = S1 (uint8), 2 = the
// Dirichlet face values (FP32, writable: the faces hold them every step).
extern "C" int gm_batch_field(gm_batch* b, int i, int field, void** ptr, int64_t* count) {
    if (!b || i < 0 || i >= b->n || !ptr || !count) return fail("gm_batch_field: bad arguments");
    const gm_grid& g = b->g[i];
    *count = (int64_t)g.n[0] * g.n[1] * g.n[2];
    if (field == 0) *ptr = b->T[b->cur][i];
    else if (field == 1) *ptr = b->S1[i];
    else if (field == 2) *ptr = b->face[i];
    else return fail("gm_batch_field: field is 0 (T), 1 (S1) or 2 (faces)");
    return 0;
}

// After T or S1 of an instance were written through gm_batch_field: carried
// k, cp of the current fields again.
extern "C" int gm_batch_refresh(gm_batch* b) {
    if (!b) return fail("gm_batch_refresh: null batch");
    for (int i = 0; i < b->n; ++i) {
        const gm_grid& g = b->g[i];
        const int64_t nn = (int64_t)g.n[0] * g.n[1] * g.n[2];
        launch_props_kc(b->dm[i], b->T[b->cur][i], b->S1[i], nn, 0, b->kc[b->cur][i], b->s);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_batch_jump_state(gm_batch* b, int i, const float* T0) {
    if (!b || !T0) return fail("gm_batch_jump_state: null argument");
    if (i < 0 || i >= b->n) return fail("gm_batch_jump_state: no such instance");
    const gm_grid& g = b->g[i];
    const int64_t nn = (int64_t)g.n[0] * g.n[1] * g.n[2];
    launch_jump_state(b->d_desc, b->dm[i], 0, T0, b->T[b->cur][i], b->S1[i], b->S1[i], 0,
                      nullptr, nullptr, nullptr, nn, b->s);
    HIPCHK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// Learned L3 operator: 3x3x3 BF16 MFMA convolutions, Adam
// ---------------------------------------------------------------------------

extern "C" int gm_conv3d_fwd(const int32_t* shape, const uint16_t* X, const uint16_t* Wk,
                             const float* bias, int relu, float* Y, uint16_t* Ybf, void* stream) {
    if (!shape || !X || !Wk || !Y) return fail("gm_conv3d_fwd: null argument");
    if (!conv_shape_ok(shape))
        return fail("gm_conv3d_fwd: shape {nb, D, H, W, cin, cout} needs positive sizes and "
                    "channel counts that are multiples of 16");
    if (((uintptr_t)X | (uintptr_t)Wk) & 7)
        return fail("gm_conv3d_fwd: X and Wk must be 8-byte aligned (4 bf16 per load)");
    launch_conv_fwd(shape, X, Wk, bias, relu, Y, Ybf, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_conv3d_fwd_ws_bytes(const int32_t* shape, int64_t* bytes) {
    if (!shape || !bytes) return fail("gm_conv3d_fwd_ws_bytes: null argument");
    if (!conv_shape_ok(shape)) return fail("gm_conv3d_fwd_ws_bytes: bad shape");
    *bytes = conv_fwd_split_bytes(shape);
    return 0;
}

extern "C" int gm_conv3d_fwd_ws(const int32_t* shape, const uint16_t* X, const uint16_t* Wk,
                                const float* bias, int relu, float* Y, uint16_t* Ybf, void* ws,
                                int64_t ws_bytes, void* stream) {
    if (!shape || !X || !Wk || !Y) return fail("gm_conv3d_fwd_ws: null argument");
    if (!conv_shape_ok(shape))
        return fail("gm_conv3d_fwd_ws: shape {nb, D, H, W, cin, cout} needs positive sizes and "
                    "channel counts that are multiples of 16");
    if (((uintptr_t)X | (uintptr_t)Wk) & 7)
        return fail("gm_conv3d_fwd_ws: X and Wk must be 8-byte aligned (4 bf16 per load)");
    if (ws && ((uintptr_t)ws & 15)) return fail("gm_conv3d_fwd_ws: ws must be 16-byte aligned");
    launch_conv_fwd_ws(shape, X, Wk, bias, relu, Y, Ybf, ws, ws_bytes, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_conv3d_bwd_data(const int32_t* shape, const uint16_t* dY, const uint16_t* Wk,
                                  uint16_t* Wt_work, float* dX, uint16_t* dXbf, void* stream) {
    if (!shape || !dY || !Wk || !Wt_work || !dX) return fail("gm_conv3d_bwd_data: null argument");
    if (!conv_shape_ok(shape)) return fail("gm_conv3d_bwd_data: bad shape");
    if (((uintptr_t)dY | (uintptr_t)Wk | (uintptr_t)Wt_work) & 7)
        return fail("gm_conv3d_bwd_data: dY, Wk and Wt_work must be 8-byte aligned");
    const int32_t sb[6] = {shape[0], shape[1], shape[2], shape[3], shape[5], shape[4]};
    launch_conv_wflip(shape[4], shape[5], Wk, Wt_work, (hipStream_t)stream);
    launch_conv_fwd(sb, dY, Wt_work, nullptr, 0, dX, dXbf, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_conv3d_bwd_weight(const int32_t* shape, const uint16_t* X, const uint16_t* dY,
                                    float* dW, void* stream) {
    if (!shape || !X || !dY || !dW) return fail("gm_conv3d_bwd_weight: null argument");
    if (!conv_shape_ok(shape)) return fail("gm_conv3d_bwd_weight: bad shape");
    if (((uintptr_t)X | (uintptr_t)dY) & 7)
        return fail("gm_conv3d_bwd_weight: X and dY must be 8-byte aligned");
    const int64_t nvox = (int64_t)shape[0] * shape[1] * shape[2] * shape[3];
    // about 1024 voxels per block, at most 2048 parts (grid z = 27 * parts);
    // 4096 for 32 x 32 channels from 2^18 voxels on (each block adds its
    // partial tile to dW with FP32 atomics: 163 us instead of 219 on 524288
    // voxels; the other kernels were 2-6% slower with longer parts)
    const bool long_parts = nvox >= ((int64_t)1 << 18) && shape[4] == 32 && shape[5] == 32;
    const int64_t per = long_parts ? 4096 : 1024;
    int64_t parts = (nvox + per - 1) / per;
    if (parts > 2048) parts = 2048;
    if (parts < 1) parts = 1;
    launch_conv_wgrad(shape, X, dY, (int)parts, dW, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_adam(float* w, uint16_t* wbf, const float* g, float* m, float* v, int64_t n,
                       float lr, float beta1, float beta2, float eps, int step, void* stream) {
    if (!w || !g || !m || !v || n < 0 || step < 1) return fail("gm_adam: bad arguments");
    if (n > 0) launch_adam(w, wbf, g, m, v, n, lr, beta1, beta2, eps, step, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_f32_to_bf16(const float* a, uint16_t* b, int64_t n, void* stream) {
    if (!a || !b || n < 0) return fail("gm_f32_to_bf16: bad arguments");
    if (n > 0) launch_f32_to_bf16(a, b, n, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

// ---- elementwise and reduction kernels of the operator's network ----------

static bool shape5_ok(const int32_t* s) {
    return s && s[0] > 0 && s[1] > 0 && s[2] > 0 && s[3] > 0 && s[4] > 0;
}

extern "C" int gm_nn_pool2(int dir, const int32_t* coarse_shape, const void* in, void* out,
                           void* stream) {
    if (!shape5_ok(coarse_shape) || !in || !out || (dir != 0 && dir != 1))
        return fail("gm_nn_pool2: bad arguments");
    const int32_t* c = coarse_shape;
    launch_nn_pool2(dir, c[0], c[1], c[2], c[3], c[4], in, out, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_nn_up2(int dir, const int32_t* coarse_shape, const void* in, void* out,
                         void* stream) {
    if (!shape5_ok(coarse_shape) || !in || !out || (dir != 0 && dir != 1))
        return fail("gm_nn_up2: bad arguments");
    const int32_t* c = coarse_shape;
    launch_nn_up2(dir, c[0], c[1], c[2], c[3], c[4], in, out, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_nn_concat(int64_t nvox, int ca, int cb, const uint16_t* A, const uint16_t* B,
                            uint16_t* Y, void* stream) {
    if (nvox < 0 || ca < 1 || cb < 1 || !A || !B || !Y) return fail("gm_nn_concat: bad arguments");
    if (nvox) launch_nn_concat(nvox, ca, cb, A, B, Y, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_nn_split(int64_t nvox, int ca, int cb, const float* dY, float* dA, float* dB,
                           void* stream) {
    if (nvox < 0 || ca < 1 || cb < 1 || !dY || !dA || !dB) return fail("gm_nn_split: bad arguments");
    if (nvox) launch_nn_split(nvox, ca, cb, dY, dA, dB, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_nn_relu_bwd(int64_t n, const float* dA, const float* dA2, const float* A,
                              float* dZ, uint16_t* dZb, void* stream) {
    if (n < 0 || !dA || !dZ) return fail("gm_nn_relu_bwd: bad arguments");
    if (n) launch_nn_relu_bwd(n, dA, dA2, A, dZ, dZb, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_nn_bias_grad(int64_t nvox, int C, const float* dZ, float* db, void* stream) {
    if (nvox < 0 || C < 1 || !dZ || !db) return fail("gm_nn_bias_grad: bad arguments");
    if (nvox) launch_nn_bias_grad(nvox, C, dZ, db, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_nn_relu_bias_bwd(int64_t nvox, int C, const float* dA, const float* dA2,
                                   const float* A, float* dZ, uint16_t* dZb, float* db,
                                   void* stream) {
    if (nvox < 0 || C < 4 || C > 256 || (C & 3) || (256 % C) || !dA || !db)
        return fail("gm_nn_relu_bias_bwd: needs dA, db and C a multiple of 4 dividing 256");
    const void* ptrs[5] = {dA, dA2, A, dZ, dZb};
    for (int i = 0; i < 5; ++i)
        if ((uintptr_t)ptrs[i] & (i == 4 ? 7 : 15))
            return fail("gm_nn_relu_bias_bwd: operands must be 16-byte (dZb 8-byte) aligned");
    if (nvox) launch_nn_relu_bias_bwd(nvox, C, dA, dA2, A, dZ, dZb, db, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_nn_mse(int64_t nvox, int C, const float* Y, const float* target, float* dY,
                         double* loss, void* stream) {
    if (nvox < 0 || C < 1 || !Y || !target || !dY || !loss) return fail("gm_nn_mse: bad arguments");
    if (nvox) launch_nn_mse(nvox, C, Y, target, dY, loss, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_nn_mse_energy(int64_t nw, int nb, int C, int nx, int ny, float lambda,
                                const float* Y, const float* target, double* wsum, float* dY,
                                double* loss, void* stream) {
    if (nw < 1 || nb < 1 || nb > 65535 || C < 1 || nx < 1 || ny < 1 ||
        nw % ((int64_t)nx * ny) || !Y || !target || !wsum || !dY || !loss)
        return fail("gm_nn_mse_energy: bad arguments");
    launch_nn_mse_energy(nw, nb, C, nx, ny, lambda, Y, target, wsum, dY, loss,
                         (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_nn_window_input(const int32_t* n_host, const float* geo_host, const float* T,
                                  const float* Tface, const float* prm_host, uint16_t* X,
                                  void* stream) {
    if (!n_host || !geo_host || !T || !prm_host || !X || n_host[0] < 1 || n_host[1] < 1 ||
        n_host[2] < 1)
        return fail("gm_nn_window_input: bad arguments");
    launch_nn_window_input(n_host, geo_host, T, Tface, prm_host, X, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_nn_window_apply(const int32_t* n_host, const float* T, const float* Tface,
                                  const float* Y, int C, float T_scale, float T_floor, float* Tn,
                                  void* stream) {
    if (!n_host || !T || !Y || !Tn || C < 1 || n_host[0] < 1 || n_host[1] < 1 || n_host[2] < 1)
        return fail("gm_nn_window_apply: bad arguments");
    launch_nn_window_apply(n_host, T, Tface, Y, C, T_scale, T_floor, Tn, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_u8_to_f32(int64_t n, const uint8_t* x, float* y, void* stream) {
    if (n < 0 || !x || !y) return fail("gm_u8_to_f32: bad arguments");
    if (n) launch_nn_u8_to_f32(n, x, y, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int gm_nn_axpy(int64_t n, float a, const float* x, float* y, void* stream) {
    if (n < 0 || !x || !y) return fail("gm_nn_axpy: bad arguments");
    if (n) launch_nn_axpy(n, a, x, y, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
}
