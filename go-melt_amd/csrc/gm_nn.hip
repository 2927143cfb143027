// Elementwise and reduction kernels of the learned L3 operator's network
// (the 3x3x3 convolutions are in gm_conv.hip).  Channels last:
// a[((b*D + z)*H + y)*W + x][C].  Activations feeding a convolution are bf16
// (uint16 bits); gradients are FP32 until they feed a convolution.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "gm_internal.h"

namespace gm {

__device__ __forceinline__ float nn_bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ uint16_t nn_f2bf(float f) {
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

static inline unsigned nn_blocks(int64_t n) { return (unsigned)((n + 255) / 256); }
static inline bool nn_al16(const void* p) { return ((uintptr_t)p & 15) == 0; }

// 2x2x2 average pooling: Y[b, z, y, x, c] = mean of X over the 8 fine voxels
// (fine extent 2D x 2H x 2W); dir 0 forward (bf16 -> bf16), dir 1 backward
// (FP32 dY -> FP32 dX = dY / 8 on each fine voxel)
__global__ void k_pool2(int nb, int D, int H, int W, int C, const uint16_t* __restrict__ X,
                        uint16_t* __restrict__ Y) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)nb * D * H * W * C;
    if (t >= n) return;
    const int c = (int)(t % C);
    int64_t v = t / C;
    const int x = (int)(v % W);
    v /= W;
    const int y = (int)(v % H);
    v /= H;
    const int z = (int)(v % D);
    const int b = (int)(v / D);
    const int W2 = 2 * W, H2 = 2 * H, D2 = 2 * D;
    float s = 0.0f;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int xx = 2 * x + (q & 1), yy = 2 * y + ((q >> 1) & 1), zz = 2 * z + (q >> 2);
        s += nn_bf2f(X[((((int64_t)b * D2 + zz) * H2 + yy) * W2 + xx) * C + c]);
    }
    Y[t] = nn_f2bf(0.125f * s);
}

// 1/8 of a gradient value, one float or four
__device__ __forceinline__ float nn_eighth(float x) { return 0.125f * x; }
__device__ __forceinline__ float4 nn_eighth(float4 x) {
    return make_float4(0.125f * x.x, 0.125f * x.y, 0.125f * x.z, 0.125f * x.w);
}

// V is float, or float4 with C counted in groups of four channels
template <typename V>
__global__ void k_pool2_bwd(int nb, int D, int H, int W, int C, const V* __restrict__ dY,
                            V* __restrict__ dX) {
    // one thread per fine element
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int W2 = 2 * W, H2 = 2 * H, D2 = 2 * D;
    const int64_t n = (int64_t)nb * D2 * H2 * W2 * C;
    if (t >= n) return;
    const int c = (int)(t % C);
    int64_t v = t / C;
    const int x = (int)(v % W2);
    v /= W2;
    const int y = (int)(v % H2);
    v /= H2;
    const int z = (int)(v % D2);
    const int b = (int)(v / D2);
    dX[t] = nn_eighth(dY[((((int64_t)b * D + z / 2) * H + y / 2) * W + x / 2) * C + c]);
}

// nearest 2x upsampling (coarse extent D, H, W): forward bf16 -> bf16 (V is
// one bf16, or uint4 = eight with C counted in groups of eight); backward
// sums the 8 fine gradients (FP32)
template <typename V>
__global__ void k_up2(int nb, int D, int H, int W, int C, const V* __restrict__ X,
                      V* __restrict__ Y) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int W2 = 2 * W, H2 = 2 * H, D2 = 2 * D;
    const int64_t n = (int64_t)nb * D2 * H2 * W2 * C;
    if (t >= n) return;
    const int c = (int)(t % C);
    int64_t v = t / C;
    const int x = (int)(v % W2);
    v /= W2;
    const int y = (int)(v % H2);
    v /= H2;
    const int z = (int)(v % D2);
    const int b = (int)(v / D2);
    Y[t] = X[((((int64_t)b * D + z / 2) * H + y / 2) * W + x / 2) * C + c];
}

__global__ void k_up2_bwd(int nb, int D, int H, int W, int C, const float* __restrict__ dY,
                          float* __restrict__ dX) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)nb * D * H * W * C;
    if (t >= n) return;
    const int c = (int)(t % C);
    int64_t v = t / C;
    const int x = (int)(v % W);
    v /= W;
    const int y = (int)(v % H);
    v /= H;
    const int z = (int)(v % D);
    const int b = (int)(v / D);
    const int W2 = 2 * W, H2 = 2 * H, D2 = 2 * D;
    float s = 0.0f;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int xx = 2 * x + (q & 1), yy = 2 * y + ((q >> 1) & 1), zz = 2 * z + (q >> 2);
        s += dY[((((int64_t)b * D2 + zz) * H2 + yy) * W2 + xx) * C + c];
    }
    dX[t] = s;
}

// channel concatenation Y = [A | B] (bf16) and its backward split (FP32);
// V is one value, or 16 bytes of them (uint4: eight bf16, float4: four
// floats) with the channel counts in those groups
template <typename V>
__global__ void k_concat(int64_t nvox, int ca, int cb, const V* __restrict__ A,
                         const V* __restrict__ B, V* __restrict__ Y) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int cc = ca + cb;
    if (t >= nvox * cc) return;
    const int64_t v = t / cc;
    const int c = (int)(t - v * cc);
    Y[t] = c < ca ? A[v * ca + c] : B[v * cb + (c - ca)];
}

template <typename V>
__global__ void k_split(int64_t nvox, int ca, int cb, const V* __restrict__ dY,
                        V* __restrict__ dA, V* __restrict__ dB) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int cc = ca + cb;
    if (t >= nvox * cc) return;
    const int64_t v = t / cc;
    const int c = (int)(t - v * cc);
    if (c < ca) dA[v * ca + c] = dY[t];
    else dB[v * cb + (c - ca)] = dY[t];
}

// ReLU backward through the layer's output A (A > 0 where the input was):
// dZ = dA (+ dA2 when given) where A > 0, else 0; FP32 and a bf16 copy
__global__ void k_relu_bwd(int64_t n, const float* __restrict__ dA, const float* __restrict__ dA2,
                           const float* __restrict__ A, float* __restrict__ dZ,
                           uint16_t* __restrict__ dZb) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float g = dA[t] + (dA2 ? dA2[t] : 0.0f);
    if (A && !(A[t] > 0.0f)) g = 0.0f;
    dZ[t] = g;
    if (dZb) dZb[t] = nn_f2bf(g);
}

// db[c] += sum over voxels of dZ[v][c]: each block sums 256 voxels of one
// 64-channel slice (blockIdx.y), one atomic per channel and block
__global__ void k_bias_grad(int64_t nvox, int C, const float* __restrict__ dZ,
                            float* __restrict__ db) {
    __shared__ float part[4][64];
    const int c = blockIdx.y * 64 + (threadIdx.x & 63);
    const int r = threadIdx.x >> 6;  // 4 row groups
    float s = 0.0f;
    const int64_t v0 = (int64_t)blockIdx.x * 256;
    if (c < C)
        for (int k = r; k < 256; k += 4) {
            const int64_t v = v0 + k;
            if (v < nvox) s += dZ[v * C + c];
        }
    part[r][threadIdx.x & 63] = s;
    __syncthreads();
    if (r == 0 && c < C) {
        const float tot = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] +
                          part[3][threadIdx.x];
        atomicAdd(db + c, tot);
    }
}

// ReLU backward fused with the bias gradient, for C a multiple of 4 that
// divides 256 and 16-byte aligned operands.  Each thread takes float4 groups
// 1024 x gridDim.x apart, so its four channels stay fixed (4 x lane mod C):
// it sums them in registers, the wave folds lanes holding the same channels
// (xor over lane bits >= log2(C / 4)), the 16 waves meet in LDS and the block
// adds one atomic per channel.  dZb = bf16(g), dZ = g when given, with
// g = (dA + dA2) where A > 0, else 0; db[c] += sum over voxels of g[v][c].
__global__ void __launch_bounds__(1024) k_relu_bias_bwd(int64_t n4, int C,
                                                        const float4* __restrict__ dA,
                                                        const float4* __restrict__ dA2,
                                                        const float4* __restrict__ A,
                                                        float4* __restrict__ dZ,
                                                        uint2* __restrict__ dZb,
                                                        float* __restrict__ db) {
    __shared__ float part[16][256];
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
    const int64_t stride = (int64_t)gridDim.x * 1024;
    for (int64_t t = (int64_t)blockIdx.x * 1024 + threadIdx.x; t < n4; t += stride) {
        const float4 a = dA[t];
        const float4 b = dA2 ? dA2[t] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float4 g = make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
        if (A) {
            const float4 m = A[t];
            if (!(m.x > 0.0f)) g.x = 0.0f;
            if (!(m.y > 0.0f)) g.y = 0.0f;
            if (!(m.z > 0.0f)) g.z = 0.0f;
            if (!(m.w > 0.0f)) g.w = 0.0f;
        }
        if (dZ) dZ[t] = g;
        if (dZb)
            dZb[t] = make_uint2((uint32_t)nn_f2bf(g.x) | ((uint32_t)nn_f2bf(g.y) << 16),
                                (uint32_t)nn_f2bf(g.z) | ((uint32_t)nn_f2bf(g.w) << 16));
        s0 += g.x;
        s1 += g.y;
        s2 += g.z;
        s3 += g.w;
    }
    const int lanes = C >> 2;  // lanes per channel period
    for (int sh = 32; sh >= lanes; sh >>= 1) {
        s0 += __shfl_xor(s0, sh, 64);
        s1 += __shfl_xor(s1, sh, 64);
        s2 += __shfl_xor(s2, sh, 64);
        s3 += __shfl_xor(s3, sh, 64);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane < lanes) {
        part[w][4 * lane + 0] = s0;
        part[w][4 * lane + 1] = s1;
        part[w][4 * lane + 2] = s2;
        part[w][4 * lane + 3] = s3;
    }
    __syncthreads();
    if ((int)threadIdx.x < C) {
        float tot = 0.0f;
#pragma unroll
        for (int k = 0; k < 16; ++k) tot += part[k][threadIdx.x];
        atomicAdd(db + threadIdx.x, tot);
    }
}

// Loss on channel 0 of the head output Y (nvox x C): 0.5 mean (Y0 - target)^2
// summed in FP64 (loss[0] += ...); dY (FP32, nvox x C) = (Y0 - target) / nvox
// in channel 0 and zero elsewhere
__global__ void __launch_bounds__(256) k_mse(int64_t nvox, int C, const float* __restrict__ Y,
                                             const float* __restrict__ target,
                                             float* __restrict__ dY, double* __restrict__ loss) {
    // a grid of at most 1024 blocks; each folds its four waves in LDS and
    // adds once (atomics on one address serialise)
    __shared__ double part[4];
    double l = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nvox; t += stride) {
        const float r = Y[t * C] - target[t];
        l += 0.5 * (double)r * (double)r / (double)nvox;
        dY[t * C] = r / (float)nvox;
        for (int c = 1; c < C; ++c) dY[t * C + c] = 0.0f;
    }
#pragma unroll
    for (int sh = 32; sh >= 1; sh >>= 1) l += __shfl_xor(l, sh, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double b = part[0] + part[1] + part[2] + part[3];
        if (b != 0.0) atomicAdd(loss, b);
    }
}

// Energy-consistent loss of nb windows of nw voxels (nvox = nb nw; nx x ny x nz
// nodes, x fastest), r = Y[., 0] - target:  L = 0.5 mean_v r_v^2 +
// 0.5 lambda mean_w m_w^2, m_w the sum of r over the window's free voxels (not
// on the Dirichlet faces W, E, S, N, bottom, whose outputs a jump does not use)
// divided by nw: the window's heat-content error, which the energy audit of a
// jump sums as C (T_new - T) over the free nodes.  Pass 1 sums r per window
// (FP64, one atomic per block and window); pass 2 writes
// dL/dY = (r_v + lambda m_w [v free]) / nvox and adds the loss.
__device__ __forceinline__ bool window_face(int64_t t, int nx, int ny, int64_t nw) {
    const int nz = (int)(nw / ((int64_t)nx * ny));
    const int ix = (int)(t % nx), iy = (int)((t / nx) % ny), iz = (int)(t / ((int64_t)nx * ny));
    (void)nz;
    return ix == 0 || ix == nx - 1 || iy == 0 || iy == ny - 1 || iz == 0;
}

__global__ void __launch_bounds__(256) k_window_rsum(int64_t nw, int nb, int C, int nx, int ny,
                                                     const float* __restrict__ Y,
                                                     const float* __restrict__ target,
                                                     double* __restrict__ wsum) {
    __shared__ double part[4];
    const int w = blockIdx.y;
    double l = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nw; t += stride) {
        const int64_t v = (int64_t)w * nw + t;
        if (!window_face(t, nx, ny, nw)) l += (double)(Y[v * C] - target[v]);
    }
#pragma unroll
    for (int sh = 32; sh >= 1; sh >>= 1) l += __shfl_xor(l, sh, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double b = part[0] + part[1] + part[2] + part[3];
        if (b != 0.0) atomicAdd(&wsum[w], b);
    }
}

__global__ void __launch_bounds__(256) k_mse_energy(int64_t nw, int nb, int C, int nx, int ny,
                                                    float lambda,
                                                    const float* __restrict__ Y,
                                                    const float* __restrict__ target,
                                                    const double* __restrict__ wsum,
                                                    float* __restrict__ dY,
                                                    double* __restrict__ loss) {
    __shared__ double part[4];
    const int w = blockIdx.y;
    const int64_t nvox = nw * nb;
    const double m = wsum[w] / (double)nw;
    double l = (blockIdx.x == 0 && threadIdx.x == 0) ? 0.5 * (double)lambda * m * m / nb : 0.0;
    const float mf = (float)m;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nw; t += stride) {
        const int64_t v = (int64_t)w * nw + t;
        const float r = Y[v * C] - target[v];
        l += 0.5 * (double)r * (double)r / (double)nvox;
        dY[v * C] = (r + (window_face(t, nx, ny, nw) ? 0.0f : lambda * mf)) / (float)nvox;
        for (int c = 1; c < C; ++c) dY[v * C + c] = 0.0f;
    }
#pragma unroll
    for (int sh = 32; sh >= 1; sh >>= 1) l += __shfl_xor(l, sh, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double b = part[0] + part[1] + part[2] + part[3];
        if (b != 0.0) atomicAdd(loss, b);
    }
}

void launch_nn_mse_energy(int64_t nw, int nb, int C, int nx, int ny, float lambda,
                          const float* Y, const float* target, double* wsum, float* dY,
                          double* loss, hipStream_t s) {
    const unsigned gx = std::min(nn_blocks(nw), 256u);
    hipLaunchKernelGGL(k_window_rsum, dim3(gx, nb), dim3(256), 0, s, nw, nb, C, nx, ny, Y,
                       target, wsum);
    hipLaunchKernelGGL(k_mse_energy, dim3(gx, nb), dim3(256), 0, s, nw, nb, C, nx, ny, lambda, Y,
                       target, wsum, dY, loss);
}

// Network input of one window (nx x ny x nz nodes, x fastest), 16 channels:
// 0: (T - T_amb) / T_scale, 1: (T_face - T_amb) / T_scale on the Dirichlet
// faces (W, E, S, N, bottom) and 0 inside, 2: the laser intensity
// exp(-3 ((x - vx)^2 + (y - vy)^2) / r^2 - 3 (z - vz)^2 / d^2), 3: P / P_scale,
// 4..15: zero.  Windows are stacked along the batch dimension.
__global__ void k_window_input(int nx, int ny, int nz, float x0, float y0, float z0, float h,
                               const float* __restrict__ T, const float* __restrict__ Tface,
                               float T_amb, float T_scale, float vx, float vy, float vz, float r,
                               float d, float P, float P_scale, uint16_t* __restrict__ X) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nn = (int64_t)nx * ny * nz;
    if (t >= nn) return;
    const int ix = (int)(t % nx), iy = (int)((t / nx) % ny), iz = (int)(t / ((int64_t)nx * ny));
    const bool face = ix == 0 || ix == nx - 1 || iy == 0 || iy == ny - 1 || iz == 0;
    const float x = x0 + ix * h, y = y0 + iy * h, z = z0 + iz * h;
    const float q = __expf(-3.0f * ((x - vx) * (x - vx) + (y - vy) * (y - vy)) / (r * r) -
                           3.0f * (z - vz) * (z - vz) / (d * d));
    uint16_t* o = X + t * 16;
    o[0] = nn_f2bf((T[t] - T_amb) / T_scale);
    o[1] = nn_f2bf(face && Tface ? (Tface[t] - T_amb) / T_scale : 0.0f);
    o[2] = nn_f2bf(q);
    o[3] = nn_f2bf(P / P_scale);
#pragma unroll
    for (int c = 4; c < 16; ++c) o[c] = 0;
}

// y += a * x (FP32)
__global__ void k_axpy(int64_t n, float a, const float* __restrict__ x, float* __restrict__ y) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] += a * x[t];
}

// T_next = T + T_scale * Y[v][0] on interior nodes; faces keep T_face
__global__ void k_window_apply(int nx, int ny, int nz, const float* __restrict__ T,
                               const float* __restrict__ Tface, const float* __restrict__ Y,
                               int C, float T_scale, float T_floor, float* __restrict__ Tn) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nn = (int64_t)nx * ny * nz;
    if (t >= nn) return;
    const int ix = (int)(t % nx), iy = (int)((t / nx) % ny), iz = (int)(t / ((int64_t)nx * ny));
    const bool face = ix == 0 || ix == nx - 1 || iy == 0 || iy == ny - 1 || iz == 0;
    float v = face ? (Tface ? Tface[t] : T[t]) : T[t] + T_scale * Y[t * C];
    Tn[t] = v < T_floor ? T_floor : v;
}

void launch_nn_pool2(int dir, int nb, int D, int H, int W, int C, const void* in, void* out,
                     hipStream_t s) {
    if (dir == 0) {
        const int64_t n = (int64_t)nb * D * H * W * C;
        hipLaunchKernelGGL(k_pool2, dim3(nn_blocks(n)), dim3(256), 0, s, nb, D, H, W, C,
                           (const uint16_t*)in, (uint16_t*)out);
    } else if (C % 4 == 0 && nn_al16(in) && nn_al16(out)) {
        const int64_t n = (int64_t)nb * 8 * D * H * W * (C / 4);
        hipLaunchKernelGGL(k_pool2_bwd<float4>, dim3(nn_blocks(n)), dim3(256), 0, s, nb, D, H, W,
                           C / 4, (const float4*)in, (float4*)out);
    } else {
        const int64_t n = (int64_t)nb * 8 * D * H * W * C;
        hipLaunchKernelGGL(k_pool2_bwd<float>, dim3(nn_blocks(n)), dim3(256), 0, s, nb, D, H, W,
                           C, (const float*)in, (float*)out);
    }
}

void launch_nn_up2(int dir, int nb, int D, int H, int W, int C, const void* in, void* out,
                   hipStream_t s) {
    if (dir == 0 && C % 8 == 0 && nn_al16(in) && nn_al16(out)) {
        const int64_t n = (int64_t)nb * 8 * D * H * W * (C / 8);
        hipLaunchKernelGGL(k_up2<uint4>, dim3(nn_blocks(n)), dim3(256), 0, s, nb, D, H, W, C / 8,
                           (const uint4*)in, (uint4*)out);
    } else if (dir == 0) {
        const int64_t n = (int64_t)nb * 8 * D * H * W * C;
        hipLaunchKernelGGL(k_up2<uint16_t>, dim3(nn_blocks(n)), dim3(256), 0, s, nb, D, H, W, C,
                           (const uint16_t*)in, (uint16_t*)out);
    } else {
        const int64_t n = (int64_t)nb * D * H * W * C;
        hipLaunchKernelGGL(k_up2_bwd, dim3(nn_blocks(n)), dim3(256), 0, s, nb, D, H, W, C,
                           (const float*)in, (float*)out);
    }
}

void launch_nn_concat(int64_t nvox, int ca, int cb, const uint16_t* A, const uint16_t* B,
                      uint16_t* Y, hipStream_t s) {
    if (ca % 8 == 0 && cb % 8 == 0 && nn_al16(A) && nn_al16(B) && nn_al16(Y))
        hipLaunchKernelGGL(k_concat<uint4>, dim3(nn_blocks(nvox * (ca + cb) / 8)), dim3(256), 0, s,
                           nvox, ca / 8, cb / 8, (const uint4*)A, (const uint4*)B, (uint4*)Y);
    else
        hipLaunchKernelGGL(k_concat<uint16_t>, dim3(nn_blocks(nvox * (ca + cb))), dim3(256), 0, s,
                           nvox, ca, cb, A, B, Y);
}

void launch_nn_split(int64_t nvox, int ca, int cb, const float* dY, float* dA, float* dB,
                     hipStream_t s) {
    if (ca % 4 == 0 && cb % 4 == 0 && nn_al16(dY) && nn_al16(dA) && nn_al16(dB))
        hipLaunchKernelGGL(k_split<float4>, dim3(nn_blocks(nvox * (ca + cb) / 4)), dim3(256), 0, s,
                           nvox, ca / 4, cb / 4, (const float4*)dY, (float4*)dA, (float4*)dB);
    else
        hipLaunchKernelGGL(k_split<float>, dim3(nn_blocks(nvox * (ca + cb))), dim3(256), 0, s,
                           nvox, ca, cb, dY, dA, dB);
}

void launch_nn_relu_bwd(int64_t n, const float* dA, const float* dA2, const float* A, float* dZ,
                        uint16_t* dZb, hipStream_t s) {
    hipLaunchKernelGGL(k_relu_bwd, dim3(nn_blocks(n)), dim3(256), 0, s, n, dA, dA2, A, dZ, dZb);
}

void launch_nn_bias_grad(int64_t nvox, int C, const float* dZ, float* db, hipStream_t s) {
    const dim3 grid((unsigned)((nvox + 255) / 256), (unsigned)((C + 63) / 64));
    hipLaunchKernelGGL(k_bias_grad, grid, dim3(256), 0, s, nvox, C, dZ, db);
}

void launch_nn_relu_bias_bwd(int64_t nvox, int C, const float* dA, const float* dA2,
                             const float* A, float* dZ, uint16_t* dZb, float* db, hipStream_t s) {
    const int64_t n4 = nvox * C / 4;
    const int64_t blocks = std::min<int64_t>((n4 + 1023) / 1024, 256);
    hipLaunchKernelGGL(k_relu_bias_bwd, dim3((unsigned)blocks), dim3(1024), 0, s, n4, C,
                       (const float4*)dA, (const float4*)dA2, (const float4*)A, (float4*)dZ,
                       (uint2*)dZb, db);
}

void launch_nn_mse(int64_t nvox, int C, const float* Y, const float* target, float* dY,
                   double* loss, hipStream_t s) {
    hipLaunchKernelGGL(k_mse, dim3(std::min(nn_blocks(nvox), 1024u)), dim3(256), 0, s, nvox, C, Y,
                       target, dY, loss);
}

void launch_nn_window_input(const int32_t* n, const float* geo, const float* T, const float* Tface,
                            const float* prm, uint16_t* X, hipStream_t s) {
    const int64_t nn = (int64_t)n[0] * n[1] * n[2];
    hipLaunchKernelGGL(k_window_input, dim3(nn_blocks(nn)), dim3(256), 0, s, n[0], n[1], n[2],
                       geo[0], geo[1], geo[2], geo[3], T, Tface, prm[0], prm[1], prm[2], prm[3],
                       prm[4], prm[5], prm[6], prm[7], prm[8], X);
}

// y = (float)x for 0/1 phase bytes
__global__ void k_u8_to_f32(int64_t n, const uint8_t* __restrict__ x, float* __restrict__ y) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] = (float)x[t];
}

void launch_nn_u8_to_f32(int64_t n, const uint8_t* x, float* y, hipStream_t s) {
    hipLaunchKernelGGL(k_u8_to_f32, dim3(nn_blocks(n)), dim3(256), 0, s, n, x, y);
}

void launch_nn_axpy(int64_t n, float a, const float* x, float* y, hipStream_t s) {
    hipLaunchKernelGGL(k_axpy, dim3(nn_blocks(n)), dim3(256), 0, s, n, a, x, y);
}

void launch_nn_window_apply(const int32_t* n, const float* T, const float* Tface, const float* Y,
                            int C, float T_scale, float T_floor, float* Tn, hipStream_t s) {
    const int64_t nn = (int64_t)n[0] * n[1] * n[2];
    hipLaunchKernelGGL(k_window_apply, dim3(nn_blocks(nn)), dim3(256), 0, s, n[0], n[1], n[2], T,
                       Tface, Y, C, T_scale, T_floor, Tn);
}

}  // namespace gm
