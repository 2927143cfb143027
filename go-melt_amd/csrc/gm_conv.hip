// 3x3x3 convolutions for the learned L3 operator: BF16 operands on the MFMA
// matrix cores (v_mfma_f32_16x16x32_bf16 in the forward kernel for inputs in
// 32-channel stages, v_mfma_f32_16x16x16_bf16 elsewhere), FP32 accumulation.
//
// Layouts (channels last):
//   activations  X[vox][C], vox = ((b*D + z)*H + y)*W + x, bf16 (uint16 bits)
//   weights      Wk[tap][Cout][Cin], tap = (dz+1)*9 + (dy+1)*3 + (dx+1), bf16
//   outputs      FP32 (and optionally a bf16 copy for the next layer)
// Stride 1, zero padding 1, so every layer keeps D x H x W.
//
// One kernel does the implicit GEMM Y[vox][o] = sum_{tap,i} X[vox+s(tap)][i]
// Wk[tap][o][i] for the forward pass and, with the flipped and transposed
// weights Wk'[tap][i][o] = Wk[26-tap][o][i], for the gradient of the input.
// The weight gradient dWk[tap][o][i] = sum_vox dY[vox][o] X[vox+s(tap)][i] is
// a second kernel that reduces over voxels (split over blocks, FP32 atomics).
//
// Tiles: a block of 4 waves computes 64 x 64 outputs; each wave a 32 x 32
// quarter as 2 x 2 MFMA tiles of 16 x 16 (k_conv_fwd32 on large volumes with
// up to 32 outputs: waves of 64 voxels x 32 outputs); K advances through LDS one tap and
// 32 input channels at a time (k_conv_fwd32, software-pipelined) or, for 16
// input channels, 16 at a time (k_conv_fwd).

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gm_internal.h"

namespace gm {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {
    // round to nearest even (finite inputs)
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

struct ConvDims {
    int nb, D, H, W;  // batch and spatial extent
    int ci, co;       // channels (multiples of 16)
    int64_t nvox;     // nb * D * H * W
};

// (x, y, z) of a voxel, decoded once per thread (32-bit divisions while the
// volume has fewer than 2^32 voxels)
struct Vox {
    int64_t v;
    int x, y, z;
};
__device__ __forceinline__ Vox vox_decode(const ConvDims& c, int64_t v) {
    Vox r;
    r.v = v;
    if (c.nvox < ((int64_t)1 << 32)) {
        const uint32_t u = (uint32_t)v, q = u / (uint32_t)c.W, q2 = q / (uint32_t)c.H;
        r.x = (int)(u - q * (uint32_t)c.W);
        r.y = (int)(q - q2 * (uint32_t)c.H);
        r.z = (int)(q2 % (uint32_t)c.D);
    } else {
        r.x = (int)(v % c.W);
        const int64_t q = v / c.W;
        r.y = (int)(q % c.H);
        r.z = (int)((q / c.H) % c.D);
    }
    return r;
}

// voxel index shifted by tap, or -1 outside the volume (zero padding)
__device__ __forceinline__ int64_t vox_shift(const ConvDims& c, const Vox& p, int tap) {
    const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
    const int xx = p.x + dx, yy = p.y + dy, zz = p.z + dz;
    if (xx < 0 || xx >= c.W || yy < 0 || yy >= c.H || zz < 0 || zz >= c.D) return -1;
    return p.v + dx + (int64_t)c.W * (dy + (int64_t)c.H * dz);
}

// Implicit-GEMM convolution.  Grid: (ceil(nvox/64), co/64 rounded up).
// LDS tiles: sA[64 vox][16 k], sB[64 out][16 k] in bf16, K-contiguous, so a
// lane's MFMA fragment (4 consecutive k) is one 8-byte LDS read.
__global__ void __launch_bounds__(256) k_conv_fwd(ConvDims c, const uint16_t* __restrict__ X,
                                                  const uint16_t* __restrict__ Wk,
                                                  const float* __restrict__ bias, int relu,
                                                  float* __restrict__ Y,
                                                  uint16_t* __restrict__ Ybf) {
    __shared__ uint16_t sA[64][16 + 4];
    __shared__ uint16_t sB[64][16 + 4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t m0 = (int64_t)blockIdx.x * 64;
    const int n0 = blockIdx.y * 64;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    v4f acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    // staging: thread t copies 4 bf16 (8 bytes) of row t/4, k (t%4)*4 .. +3
    const int srow = tid >> 2, sk = (tid & 3) * 4;
    const int64_t vrow = m0 + srow;
    const Vox pv = vox_decode(c, vrow < c.nvox ? vrow : 0);
    const int kchunks = c.ci / 16;
    for (int tap = 0; tap < 27; ++tap) {
        const int64_t vs = vrow < c.nvox ? vox_shift(c, pv, tap) : -1;
        for (int kc = 0; kc < kchunks; ++kc) {
            const int i0 = kc * 16 + sk;
            uint2 av = make_uint2(0u, 0u), bv = make_uint2(0u, 0u);
            if (vs >= 0) av = *(const uint2*)(X + vs * c.ci + i0);
            const int o = n0 + srow;
            if (o < c.co) bv = *(const uint2*)(Wk + ((int64_t)tap * c.co + o) * c.ci + i0);
            __syncthreads();
            *(uint2*)&sA[srow][sk] = av;
            *(uint2*)&sB[srow][sk] = bv;
            __syncthreads();
            // fragments: A[i][k] at lane i + 16 (k/4), item k%4; B[k][j] likewise
            const int fi = lane & 15, fk = (lane >> 4) * 4;
            v4s fa[2], fb[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const uint2 ua = *(const uint2*)&sA[wm + t * 16 + fi][fk];
                const uint2 ub = *(const uint2*)&sB[wn + t * 16 + fi][fk];
                fa[t] = __builtin_bit_cast(v4s, ua);
                fb[t] = __builtin_bit_cast(v4s, ub);
            }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fa[a], fb[b], acc[a][b],
                                                                          0, 0, 0);
        }
    }
    // D[i][j] at lane j + 16 (i/4), item i%4
    const int dj = lane & 15, di = (lane >> 4) * 4;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t v = m0 + wm + a * 16 + di + r;
                const int o = n0 + wn + b * 16 + dj;
                if (v < c.nvox && o < c.co) {
                    float y = acc[a][b][r] + (bias ? bias[o] : 0.0f);
                    if (relu && y < 0.0f) y = 0.0f;
                    Y[v * c.co + o] = y;
                    if (Ybf) Ybf[v * c.co + o] = f2bf(y);
                }
            }
}

// The same convolution for cin a multiple of 32, software-pipelined: a K
// stage is G sub-stages of one tap x 32 input channels (one
// v_mfma_f32_16x16x32_bf16 k-step each), each thread stages 16 bytes of A and
// of B per sub-stage and row pass, and the next stage's global loads are in
// flight while the current stage computes.  Sub-stages past the last tap are
// zero.  CI16: cin = 16, and a sub-stage is two taps x 16 channels (14
// sub-stages, the 28th tap zero).
//
// Tiles: BM voxels x BN outputs per block, four waves of WM x WN, each wave
// (WM/16) x (WN/16) MFMA tiles of 16 x 16.  Larger wave tiles read fewer LDS
// bytes per MFMA (a 64 x 64 wave tile reads 128 bytes per lane for 16 MFMAs,
// a 32 x 32 one 64 bytes for 4).  Every output's K sum runs through the same
// stages in the same order whatever the tile shape, so the shapes give the
// same bits.
//
// Split K (part != nullptr): block z of gridDim.z runs stages z*spb ..
// z*spb + spb - 1 and writes its raw sums to part[z][vox][out]; the output
// is then k_conv_part_sum's fixed-order sum of the parts.
template <int BM, int BN, bool CI16 = false, int G = 1, int WM = 32, int WN = 32>
__global__ void __launch_bounds__(256) k_conv_fwd32(ConvDims c, const uint16_t* __restrict__ X,
                                                    const uint16_t* __restrict__ Wk,
                                                    const float* __restrict__ bias, int relu,
                                                    float* __restrict__ Y,
                                                    uint16_t* __restrict__ Ybf,
                                                    float* __restrict__ part, int spb) {
    static_assert((BM / WM) * (BN / WN) == 4, "four waves");
    static_assert(BM % 64 == 0 && WM % 16 == 0 && WN % 16 == 0, "tile shape");
    constexpr int NA = BM / 64;             // A rows per thread (row passes)
    constexpr int NB = BN < 64 ? 1 : BN / 64;  // B row passes
    constexpr int TM = WM / 16, TN = WN / 16;
    constexpr int KW = G * 32 + 8;
    __shared__ uint16_t sA[BM][KW];
    __shared__ uint16_t sB[BN][KW];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // XCD-aware tiles: a run of consecutive voxel tiles (with all their output
    // tiles) per XCD, so the tap-shifted rows of neighbouring tiles hit its L2
    const int tile = xcd_tile(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int mt = tile / gridDim.y;
    const int64_t m0 = (int64_t)mt * BM;
    const int n0 = (tile - mt * gridDim.y) * BN;
    const int wm = (wave % (BM / WM)) * WM, wn = (wave / (BM / WM)) * WN;
    v4f acc[TM][TN];
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b) acc[a][b] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    // staging: 16 bytes (8 bf16) of row r, k (t%4)*8 .. +7, rows r = t/4 + 64 q
    const int srow = tid >> 2, sk = (tid & 3) * 8;
    Vox pv[NA];
    int64_t vrow[NA];
#pragma unroll
    for (int q = 0; q < NA; ++q) {
        vrow[q] = m0 + srow + 64 * q;
        pv[q] = vox_decode(c, vrow[q] < c.nvox ? vrow[q] : 0);
    }
    const int kch = CI16 ? 1 : c.ci / 32;
    const int nsub = CI16 ? 14 : 27 * kch;
    const int nst = (nsub + G - 1) / G;
    const int st0 = (int)blockIdx.z * spb;
    const int st1 = st0 + spb < nst ? st0 + spb : nst;
    auto fetch = [&](int st, uint4 (*av)[NA], uint4 (*bv)[NB]) {
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int sub = st * G + g;
            int tap, i0;
            if constexpr (CI16) {
                tap = 2 * sub + (sk >> 4);  // k 0-15: tap 2 sub, k 16-31: tap 2 sub + 1
                i0 = sk & 15;
            } else {
                tap = sub / kch;
                i0 = (sub - tap * kch) * 32 + sk;
            }
            const bool tv = sub < nsub && tap < 27;
#pragma unroll
            for (int q = 0; q < NA; ++q) {
                const int64_t vs = vrow[q] < c.nvox && tv ? vox_shift(c, pv[q], tap) : -1;
                av[g][q] = make_uint4(0u, 0u, 0u, 0u);
                if (vs >= 0) av[g][q] = *(const uint4*)(X + vs * c.ci + i0);
            }
#pragma unroll
            for (int q = 0; q < NB; ++q) {
                const int br = srow + 64 * q, o = n0 + br;  // B row (output)
                bv[g][q] = make_uint4(0u, 0u, 0u, 0u);
                if (br < BN && o < c.co && tv)
                    bv[g][q] = *(const uint4*)(Wk + ((int64_t)tap * c.co + o) * c.ci + i0);
            }
        }
    };
    uint4 av[G][NA], bv[G][NB];
    fetch(st0, av, bv);
    const int fi = lane & 15, fk8 = (lane >> 4) * 8;
    for (int st = st0; st < st1; ++st) {
        __syncthreads();
#pragma unroll
        for (int g = 0; g < G; ++g) {
#pragma unroll
            for (int q = 0; q < NA; ++q) *(uint4*)&sA[srow + 64 * q][g * 32 + sk] = av[g][q];
#pragma unroll
            for (int q = 0; q < NB; ++q)
                if (srow + 64 * q < BN) *(uint4*)&sB[srow + 64 * q][g * 32 + sk] = bv[g][q];
        }
        __syncthreads();
        if (st + 1 < st1) fetch(st + 1, av, bv);
        // each sub-stage's 32 channels in one K = 32 MFMA: each lane gives 8
        // consecutive channels of its row (A) and column (B), the same k
        // positions for both operands
#pragma unroll
        for (int g = 0; g < G; ++g) {
            v8bf fa[TM], fb[TN];
#pragma unroll
            for (int t = 0; t < TM; ++t)
                fa[t] = __builtin_bit_cast(v8bf,
                                           *(const uint4*)&sA[wm + t * 16 + fi][g * 32 + fk8]);
#pragma unroll
            for (int t = 0; t < TN; ++t)
                fb[t] = __builtin_bit_cast(v8bf,
                                           *(const uint4*)&sB[wn + t * 16 + fi][g * 32 + fk8]);
#pragma unroll
            for (int a = 0; a < TM; ++a)
#pragma unroll
                for (int b = 0; b < TN; ++b)
                    acc[a][b] =
                        __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[a], fb[b], acc[a][b], 0, 0, 0);
        }
    }
    const int dj = lane & 15, di = (lane >> 4) * 4;
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t v = m0 + wm + a * 16 + di + r;
                const int oo = n0 + wn + b * 16 + dj;
                if (v < c.nvox && oo < c.co) {
                    if (part) {
                        part[(int64_t)blockIdx.z * c.nvox * c.co + v * c.co + oo] = acc[a][b][r];
                        continue;
                    }
                    float y = acc[a][b][r] + (bias ? bias[oo] : 0.0f);
                    if (relu && y < 0.0f) y = 0.0f;
                    Y[v * c.co + oo] = y;
                    if (Ybf) Ybf[v * c.co + oo] = f2bf(y);
                }
            }
}

// Y = sum over the S parts (in part order) + bias, relu, bf16 copy
__global__ void k_conv_part_sum(const float* __restrict__ part, int S, int64_t n, int co,
                                const float* __restrict__ bias, int relu, float* __restrict__ Y,
                                uint16_t* __restrict__ Ybf) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float y = 0.0f;
    for (int z = 0; z < S; ++z) y += part[(int64_t)z * n + t];
    if (bias) y += bias[t % co];
    if (relu && y < 0.0f) y = 0.0f;
    Y[t] = y;
    if (Ybf) Ybf[t] = f2bf(y);
}

// Launch plan of the pipelined forward kernel: tile shape, sub-stages per
// barrier pair and the K split.
struct FwdPlan {
    bool fwd32 = false, ci16 = false, small = false, wide = false;
    int bm = 64, bn = 64;
    dim3 grid;
    int nst = 0;     // stages of the whole K
    int splits = 1;  // K splits (grid z)
    int spb = 1;     // stages per split
};

static FwdPlan plan_fwd(const ConvDims& c, bool al16) {
    FwdPlan p;
    p.grid = dim3((unsigned)((c.nvox + 63) / 64), (unsigned)((c.co + 63) / 64));
    if (!al16 || !(c.ci % 32 == 0 || c.ci == 16)) return p;
    p.fwd32 = true;
    p.ci16 = c.ci == 16;
    // the pipelined kernel reads 16 bytes at a time; up to 32 outputs it
    // takes 128 voxels x 32 outputs per block (no idle half of the tile)
    if (c.co <= 32) {
        p.bm = 128;
        p.bn = 32;
        p.grid = dim3((unsigned)((c.nvox + 127) / 128), 1);
    }
    // Few voxels (below 2^18; one window is 2^17 at full resolution) leave
    // the chip with few resident blocks; three sub-stages per barrier pair
    // then hide more of each stage's latency (batch-1 forward 0.86 -> 0.73
    // ms).  With many voxels one sub-stage is faster (more resident blocks).
    p.small = c.nvox < ((int64_t)1 << 18);
    // Many voxels and up to 32 outputs: blocks of 256 voxels x 32 outputs,
    // each wave 64 x 32 (fewer LDS bytes per MFMA; 8-11% faster on 2^21
    // voxels).  Wider outputs keep 64 x 64 blocks of 32 x 32 waves: 256 x 64
    // and 128 x 128 blocks of 64 x 64 waves measured 5-10% slower and the
    // same there.
    p.wide = !p.small && c.co <= 32;
    if (p.wide) {
        p.bm = 256;
        p.bn = 32;
        p.grid = dim3((unsigned)((c.nvox + 255) / 256), 1);
    }
    const int G = p.small ? 3 : 1;
    const int nsub = p.ci16 ? 14 : 27 * (c.ci / 32);
    p.nst = (nsub + G - 1) / G;
    p.spb = p.nst;
    // Coarse volumes give fewer blocks than the chip has CUs: split K over
    // up to 16 blocks per tile, aiming at ~512 blocks, each with at least
    // two stages.
    const int64_t blocks = (int64_t)p.grid.x * p.grid.y;
    if (blocks < 128) {
        int S = (int)(512 / blocks);
        if (S > 16) S = 16;
        if (S > p.nst / 2) S = p.nst / 2;
        if (S >= 2) {
            p.spb = (p.nst + S - 1) / S;
            p.splits = (p.nst + p.spb - 1) / p.spb;
        }
    }
    return p;
}

template <int BM, int BN, bool CI16>
static void launch_fwd32(const FwdPlan& p, const ConvDims& c, const uint16_t* X,
                         const uint16_t* Wk, const float* bias, int relu, float* Y,
                         uint16_t* Ybf, float* part, hipStream_t s) {
    const dim3 g(p.grid.x, p.grid.y, part ? p.splits : 1);
    const int spb = part ? p.spb : p.nst;
    if (p.small)
        hipLaunchKernelGGL((k_conv_fwd32<BM, BN, CI16, 3>), g, dim3(256), 0, s, c, X, Wk, bias,
                           relu, Y, Ybf, part, spb);
    else
        hipLaunchKernelGGL((k_conv_fwd32<BM, BN, CI16, 1>), g, dim3(256), 0, s, c, X, Wk, bias,
                           relu, Y, Ybf, part, spb);
}

// many voxels, up to 32 outputs: 256 x 32 blocks of 64 x 32 waves, one
// sub-stage per barrier pair
static void launch_fwd32_wide(const FwdPlan& p, const ConvDims& c, const uint16_t* X,
                              const uint16_t* Wk, const float* bias, int relu, float* Y,
                              uint16_t* Ybf, float* part, bool ci16, hipStream_t s) {
    const dim3 g(p.grid.x, p.grid.y, part ? p.splits : 1);
    const int spb = part ? p.spb : p.nst;
    if (ci16)
        hipLaunchKernelGGL((k_conv_fwd32<256, 32, true, 1, 64, 32>), g, dim3(256), 0, s, c, X,
                           Wk, bias, relu, Y, Ybf, part, spb);
    else
        hipLaunchKernelGGL((k_conv_fwd32<256, 32, false, 1, 64, 32>), g, dim3(256), 0, s, c, X,
                           Wk, bias, relu, Y, Ybf, part, spb);
}

// Wt[tap][i][o] = Wk[26 - tap][o][i] (weights of the input-gradient pass)
__global__ void k_conv_wflip(int ci, int co, const uint16_t* __restrict__ Wk,
                             uint16_t* __restrict__ Wt) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)27 * ci * co;
    if (t >= n) return;
    const int tap = (int)(t / ((int64_t)ci * co));
    const int rem = (int)(t - (int64_t)tap * ci * co);
    const int i = rem / co, o = rem - i * co;
    Wt[t] = Wk[((int64_t)(26 - tap) * co + o) * ci + i];
}

// Weight gradient: dW[tap][o][i] += sum over this block's voxels of
// dY[v][o] X[v + s(tap)][i].  Grid: (co/64 rounded up, ci/64 rounded up,
// 27 * vsplit); each block reduces nvox/vsplit voxels in chunks of 16 and
// adds its 64 x 64 tile with FP32 atomics (dW zeroed by the caller).
__global__ void __launch_bounds__(256) k_conv_wgrad(ConvDims c, const uint16_t* __restrict__ X,
                                                    const uint16_t* __restrict__ dY, int vsplit,
                                                    float* __restrict__ dW) {
    __shared__ uint16_t sA[64][16 + 4];  // [o][vox]
    __shared__ uint16_t sB[64][16 + 4];  // [i][vox]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int o0 = blockIdx.x * 64, i0 = blockIdx.y * 64;
    const int tap = blockIdx.z % 27, part = blockIdx.z / 27;
    const int64_t per = (c.nvox + vsplit - 1) / vsplit;
    const int64_t vbeg = (int64_t)part * per;
    const int64_t vend = vbeg + per < c.nvox ? vbeg + per : c.nvox;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    v4f acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    // staging: thread t loads 4 channels (8 bytes) of voxel t/16, channels
    // (t%16)*4 .. +3 of the 64-wide tile, and writes them transposed
    const int sv = tid >> 4, sc = (tid & 15) * 4;
    for (int64_t vb = vbeg; vb < vend; vb += 16) {
        const int64_t v = vb + sv;
        uint2 dv = make_uint2(0u, 0u), xv = make_uint2(0u, 0u);
        if (v < vend) {
            if (o0 + sc < c.co) dv = *(const uint2*)(dY + v * c.co + o0 + sc);
            const int64_t vs = vox_shift(c, vox_decode(c, v), tap);
            if (vs >= 0 && i0 + sc < c.ci) xv = *(const uint2*)(X + vs * c.ci + i0 + sc);
        }
        __syncthreads();
        const uint16_t* d4 = (const uint16_t*)&dv;
        const uint16_t* x4 = (const uint16_t*)&xv;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            sA[sc + q][sv] = d4[q];
            sB[sc + q][sv] = x4[q];
        }
        __syncthreads();
        const int fi = lane & 15, fk = (lane >> 4) * 4;
        v4s fa[2], fb[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            fa[t] = __builtin_bit_cast(v4s, *(const uint2*)&sA[wm + t * 16 + fi][fk]);
            fb[t] = __builtin_bit_cast(v4s, *(const uint2*)&sB[wn + t * 16 + fi][fk]);
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
                acc[a][b] =
                    __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fa[a], fb[b], acc[a][b], 0, 0, 0);
    }
    const int dj = lane & 15, di = (lane >> 4) * 4;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int o = o0 + wm + a * 16 + di + r;
                const int i = i0 + wn + b * 16 + dj;
                if (o < c.co && i < c.ci)
                    atomicAdd(dW + ((int64_t)tap * c.co + o) * c.ci + i, acc[a][b][r]);
            }
}

// Weight gradient for cin and cout multiples of 32 (8 channels per 16-byte
// load): stages of 32 voxels staged voxel-major ([vox][channel], 16-byte LDS
// writes, no transpose), the next stage's loads in flight during the MFMAs;
// a lane's fragment (4 consecutive voxels of one channel) is 4 LDS reads.
// Tiles of OT outputs x 4096 / OT inputs: 64 x 64 (waves 2 x 2), or 32 x 128
// (waves along the inputs) so that a 32-output layer fills whole tiles.
template <int OT>
__global__ void __launch_bounds__(256) k_conv_wgrad32(ConvDims c, const uint16_t* __restrict__ X,
                                                      const uint16_t* __restrict__ dY, int vsplit,
                                                      float* __restrict__ dW) {
    constexpr int IT = 4096 / OT;  // inputs per tile
    constexpr int NX = IT / 64;    // X loads per thread and stage
    __shared__ uint16_t sA[32][OT + 8];  // [vox][o]
    __shared__ uint16_t sB[32][IT + 8];  // [vox][i]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int o0 = blockIdx.x * OT, i0 = blockIdx.y * IT;
    const int tap = blockIdx.z % 27, part = blockIdx.z / 27;
    const int64_t per = (c.nvox + vsplit - 1) / vsplit;
    const int64_t vbeg = (int64_t)part * per;
    const int64_t vend = vbeg + per < c.nvox ? vbeg + per : c.nvox;
    const int wm = OT == 64 ? (wave & 1) * 32 : 0, wn = OT == 64 ? (wave >> 1) * 32 : wave * 32;
    v4f acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    // staging: 16-byte chunks (8 channels) of the stage's 32 voxels, chunk q
    // = voxel q / (T/8), channels (q % (T/8)) * 8 of the T-wide tile: dY has
    // 4 OT chunks (threads below 4 OT), the shifted X 4 IT (NX per thread)
    const int sva = tid / (OT / 8), sca = (tid % (OT / 8)) * 8;
    auto fetch = [&](int64_t vb, uint4& dv, uint4* xv) {
        dv = make_uint4(0u, 0u, 0u, 0u);
        if (tid < 4 * OT && vb + sva < vend && o0 + sca < c.co)
            dv = *(const uint4*)(dY + (vb + sva) * c.co + o0 + sca);
#pragma unroll
        for (int k = 0; k < NX; ++k) {
            const int q = tid + 256 * k, svb = q / (IT / 8), scb = (q % (IT / 8)) * 8;
            const int64_t v = vb + svb;
            xv[k] = make_uint4(0u, 0u, 0u, 0u);
            if (v < vend && i0 + scb < c.ci) {
                const int64_t vs = vox_shift(c, vox_decode(c, v), tap);
                if (vs >= 0) xv[k] = *(const uint4*)(X + vs * c.ci + i0 + scb);
            }
        }
    };
    uint4 dv, xv[NX];
    fetch(vbeg, dv, xv);
    const int fi = lane & 15, fk = (lane >> 4) * 4;
    for (int64_t vb = vbeg; vb < vend; vb += 32) {
        __syncthreads();
        if (tid < 4 * OT) *(uint4*)&sA[sva][sca] = dv;
#pragma unroll
        for (int k = 0; k < NX; ++k) {
            const int q = tid + 256 * k;
            *(uint4*)&sB[q / (IT / 8)][(q % (IT / 8)) * 8] = xv[k];
        }
        __syncthreads();
        if (vb + 32 < vend) fetch(vb + 32, dv, xv);
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            v4s fa[2], fb[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int ro = wm + t * 16 + fi, ri = wn + t * 16 + fi, k0 = ks * 16 + fk;
                fa[t] = v4s{(short)sA[k0][ro], (short)sA[k0 + 1][ro], (short)sA[k0 + 2][ro],
                            (short)sA[k0 + 3][ro]};
                fb[t] = v4s{(short)sB[k0][ri], (short)sB[k0 + 1][ri], (short)sB[k0 + 2][ri],
                            (short)sB[k0 + 3][ri]};
            }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fa[a], fb[b], acc[a][b],
                                                                          0, 0, 0);
        }
    }
    const int dj = lane & 15, di = (lane >> 4) * 4;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int o = o0 + wm + a * 16 + di + r;
                const int i = i0 + wn + b * 16 + dj;
                if (o < c.co && i < c.ci)
                    atomicAdd(dW + ((int64_t)tap * c.co + o) * c.ci + i, acc[a][b][r]);
            }
}

// Weight gradient for cout = CO and cin = CI, each 16 or 32 (the
// full-resolution layers): each wave computes the whole CO x CI tile of one
// tap over its own quarter of a 128-voxel stage (no idle half tiles), and
// adds it with FP32 atomics.
template <int CO, int CI>
__global__ void __launch_bounds__(256) k_conv_wgrad32x(ConvDims c, const uint16_t* __restrict__ X,
                                                       const uint16_t* __restrict__ dY,
                                                       int vsplit, float* __restrict__ dW) {
    constexpr int NA = CO / 16;           // MFMA tiles along cout
    constexpr int NB = CI / 16;           // MFMA tiles along cin
    __shared__ uint16_t sA[128][CO + 8];  // [vox][o]
    __shared__ uint16_t sB[128][CI + 8];  // [vox][i]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tap = blockIdx.z % 27, part = blockIdx.z / 27;
    const int64_t per = (c.nvox + vsplit - 1) / vsplit;
    const int64_t vbeg = (int64_t)part * per;
    const int64_t vend = vbeg + per < c.nvox ? vbeg + per : c.nvox;
    v4f acc[NA][NB];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[a][b] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    // staging: 16 bytes (8 channels) of voxels t/4 and t/4 + 64, channels (t%4)*8
    const int sv = tid >> 2, sc = (tid & 3) * 8;
    auto fetch = [&](int64_t vb, uint4* dv, uint4* xv) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int64_t v = vb + sv + 64 * q;
            dv[q] = make_uint4(0u, 0u, 0u, 0u);
            xv[q] = make_uint4(0u, 0u, 0u, 0u);
            if (v < vend) {
                if (sc < CO) dv[q] = *(const uint4*)(dY + v * CO + sc);
                const int64_t vs = vox_shift(c, vox_decode(c, v), tap);
                if (vs >= 0 && sc < CI) xv[q] = *(const uint4*)(X + vs * CI + sc);
            }
        }
    };
    uint4 dv[2], xv[2];
    fetch(vbeg, dv, xv);
    const int fi = lane & 15, fk = (lane >> 4) * 4;
    const int kw = wave * 32;  // this wave's voxels of the stage
    for (int64_t vb = vbeg; vb < vend; vb += 128) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            if (sc < CO) *(uint4*)&sA[sv + 64 * q][sc] = dv[q];
            if (sc < CI) *(uint4*)&sB[sv + 64 * q][sc] = xv[q];
        }
        __syncthreads();
        if (vb + 128 < vend) fetch(vb + 128, dv, xv);
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const int k0 = kw + ks * 16 + fk;
            v4s fa[NA], fb[NB];
#pragma unroll
            for (int t = 0; t < NA; ++t) {
                const int r = t * 16 + fi;
                fa[t] = v4s{(short)sA[k0][r], (short)sA[k0 + 1][r], (short)sA[k0 + 2][r],
                            (short)sA[k0 + 3][r]};
            }
#pragma unroll
            for (int t = 0; t < NB; ++t) {
                const int r = t * 16 + fi;
                fb[t] = v4s{(short)sB[k0][r], (short)sB[k0 + 1][r], (short)sB[k0 + 2][r],
                            (short)sB[k0 + 3][r]};
            }
#pragma unroll
            for (int a = 0; a < NA; ++a)
#pragma unroll
                for (int b = 0; b < NB; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fa[a], fb[b], acc[a][b],
                                                                          0, 0, 0);
        }
    }
    const int dj = lane & 15, di = (lane >> 4) * 4;
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                atomicAdd(dW + ((int64_t)tap * CO + a * 16 + di + r) * CI + b * 16 + dj,
                          acc[a][b][r]);
}

// Adam on FP32 master weights, with the bf16 copy the kernels read
__global__ void k_adam(float* __restrict__ w, uint16_t* __restrict__ wbf,
                       const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gi = g[t];
    const float mi = b1 * m[t] + (1.0f - b1) * gi;
    const float vi = b2 * v[t] + (1.0f - b2) * gi * gi;
    m[t] = mi;
    v[t] = vi;
    const float wi = w[t] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    w[t] = wi;
    if (wbf) wbf[t] = f2bf(wi);
}

__global__ void k_f32_to_bf16(const float* __restrict__ a, uint16_t* __restrict__ b, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) b[t] = f2bf(a[t]);
}

static ConvDims dims(const int32_t* s) {
    ConvDims c;
    c.nb = s[0];
    c.D = s[1];
    c.H = s[2];
    c.W = s[3];
    c.ci = s[4];
    c.co = s[5];
    c.nvox = (int64_t)c.nb * c.D * c.H * c.W;
    return c;
}

bool conv_shape_ok(const int32_t* s) {
    for (int q = 0; q < 6; ++q)
        if (s[q] < 1) return false;
    return s[4] % 16 == 0 && s[5] % 16 == 0 &&
           (int64_t)s[0] * s[1] * s[2] * s[3] < ((int64_t)1 << 40);
}

static void launch_plan(const FwdPlan& p, const ConvDims& c, const uint16_t* X,
                        const uint16_t* Wk, const float* bias, int relu, float* Y, uint16_t* Ybf,
                        float* part, hipStream_t s) {
    if (!p.fwd32) {
        hipLaunchKernelGGL(k_conv_fwd, p.grid, dim3(256), 0, s, c, X, Wk, bias, relu, Y, Ybf);
        return;
    }
    if (p.wide) {
        if (p.ci16) launch_fwd32_wide(p, c, X, Wk, bias, relu, Y, Ybf, part, true, s);
        else launch_fwd32_wide(p, c, X, Wk, bias, relu, Y, Ybf, part, false, s);
    } else if (p.bm == 128) {
        if (p.ci16) launch_fwd32<128, 32, true>(p, c, X, Wk, bias, relu, Y, Ybf, part, s);
        else launch_fwd32<128, 32, false>(p, c, X, Wk, bias, relu, Y, Ybf, part, s);
    } else {
        if (p.ci16) launch_fwd32<64, 64, true>(p, c, X, Wk, bias, relu, Y, Ybf, part, s);
        else launch_fwd32<64, 64, false>(p, c, X, Wk, bias, relu, Y, Ybf, part, s);
    }
    if (part) {
        const int64_t n = c.nvox * c.co;
        hipLaunchKernelGGL(k_conv_part_sum, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                           part, p.splits, n, c.co, bias, relu, Y, Ybf);
    }
}

void launch_conv_fwd(const int32_t* shape, const uint16_t* X, const uint16_t* Wk,
                     const float* bias, int relu, float* Y, uint16_t* Ybf, hipStream_t s) {
    const ConvDims c = dims(shape);
    const bool al16 = (((uintptr_t)X | (uintptr_t)Wk) & 15) == 0;
    launch_plan(plan_fwd(c, al16), c, X, Wk, bias, relu, Y, Ybf, nullptr, s);
}

int64_t conv_fwd_split_bytes(const int32_t* shape) {
    const ConvDims c = dims(shape);
    const FwdPlan p = plan_fwd(c, true);
    return p.fwd32 && p.splits > 1 ? (int64_t)p.splits * c.nvox * c.co * 4 : 0;
}

void launch_conv_fwd_ws(const int32_t* shape, const uint16_t* X, const uint16_t* Wk,
                        const float* bias, int relu, float* Y, uint16_t* Ybf, void* ws,
                        int64_t ws_bytes, hipStream_t s) {
    const ConvDims c = dims(shape);
    const bool al16 = (((uintptr_t)X | (uintptr_t)Wk) & 15) == 0;
    const FwdPlan p = plan_fwd(c, al16);
    const bool split = p.fwd32 && p.splits > 1 && ws &&
                       ws_bytes >= (int64_t)p.splits * c.nvox * c.co * 4;
    launch_plan(p, c, X, Wk, bias, relu, Y, Ybf, split ? (float*)ws : nullptr, s);
}

void launch_conv_wflip(int ci, int co, const uint16_t* Wk, uint16_t* Wt, hipStream_t s) {
    const int64_t n = (int64_t)27 * ci * co;
    hipLaunchKernelGGL(k_conv_wflip, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ci, co,
                       Wk, Wt);
}

void launch_conv_wgrad(const int32_t* shape, const uint16_t* X, const uint16_t* dY, int vsplit,
                       float* dW, hipStream_t s) {
    const ConvDims c = dims(shape);
    const dim3 grid((unsigned)((c.co + 63) / 64), (unsigned)((c.ci + 63) / 64),
                    (unsigned)(27 * vsplit));
    const bool al16 = (((uintptr_t)X | (uintptr_t)dY) & 15) == 0;
    if ((c.co == 32 || c.co == 16) && (c.ci == 32 || c.ci == 16) && al16) {
        const dim3 g1(1, 1, (unsigned)(27 * vsplit));
        if (c.co == 32 && c.ci == 32)
            hipLaunchKernelGGL((k_conv_wgrad32x<32, 32>), g1, dim3(256), 0, s, c, X, dY, vsplit, dW);
        else if (c.co == 32)
            hipLaunchKernelGGL((k_conv_wgrad32x<32, 16>), g1, dim3(256), 0, s, c, X, dY, vsplit, dW);
        else if (c.ci == 32)
            hipLaunchKernelGGL((k_conv_wgrad32x<16, 32>), g1, dim3(256), 0, s, c, X, dY, vsplit, dW);
        else
            hipLaunchKernelGGL((k_conv_wgrad32x<16, 16>), g1, dim3(256), 0, s, c, X, dY, vsplit, dW);
    }
    else if (c.ci % 32 == 0 && c.co == 32 && al16) {
        const dim3 g2(1, (unsigned)((c.ci + 127) / 128), (unsigned)(27 * vsplit));
        hipLaunchKernelGGL(k_conv_wgrad32<32>, g2, dim3(256), 0, s, c, X, dY, vsplit, dW);
    } else if (c.ci % 32 == 0 && c.co % 32 == 0 && al16)
        hipLaunchKernelGGL(k_conv_wgrad32<64>, grid, dim3(256), 0, s, c, X, dY, vsplit, dW);
    else
        hipLaunchKernelGGL(k_conv_wgrad, grid, dim3(256), 0, s, c, X, dY, vsplit, dW);
}

void launch_adam(float* w, uint16_t* wbf, const float* g, float* m, float* v, int64_t n, float lr,
                 float b1, float b2, float eps, int step, hipStream_t s) {
    const float c1 = 1.0f / (1.0f - powf(b1, (float)step));
    const float c2 = 1.0f / (1.0f - powf(b2, (float)step));
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, w, wbf, g, m,
                       v, n, lr, b1, b2, eps, c1, c2);
}

void launch_f32_to_bf16(const float* a, uint16_t* b, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_f32_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, b,
                       n);
}

}  // namespace gm
