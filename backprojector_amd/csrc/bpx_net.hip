// Octree proposal net on gfx950 MFMA: 3x3x3 convolutions over a level grid, 16 channels,
// channels-last BF16 activations, FP32 accumulation (v_mfma_f32_16x16x16_bf16).
//
// Grid voxel v = (x*n1 + y)*n2 + z (z fastest), activation element (v, c) at v*16 + c.
// Weights W[t][cout][cin] (t = (dx+1)*9 + (dy+1)*3 + (dz+1)), bias[cout].
//
// MFMA 16x16x16 operand layout used here (lane l = 0..63, q = l / 16, r = l % 16):
//   A (M x K): lane holds A[r][4q .. 4q+3];  B (K x N): lane holds B[4q .. 4q+3][r];
//   D (M x N): lane holds D[4q .. 4q+3][r].
// Forward: one wave computes 16 consecutive z positions (M) x 16 output channels (N), K loops
// over the 27 taps x 16 input channels.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "bpx_internal.hpp"

namespace bpx {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t to_bf16(float x) {
    const uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// ------------------------------------------------------------------ forward / data gradient
// act: 0 = none, 1 = ReLU.  Writes bf16 (yb) and/or fp32 (yf) outputs for couts < cout_valid.
// The work-group copies the layer's weights (27 x 16 x 16 bf16, 13.5 KB) to LDS once; each
// wave then issues its 27 input loads before the 27 MFMAs (same taps, same order, same
// operands as a tap-by-tap loop: the sums are unchanged).
__global__ __launch_bounds__(256) void conv3_fwd_kernel(const uint16_t* __restrict__ X,
                                                        const uint16_t* __restrict__ W,
                                                        const float* __restrict__ bias,
                                                        uint16_t* __restrict__ yb, float* __restrict__ yf,
                                                        uint32_t n0, uint32_t n1, uint32_t n2,
                                                        int act, uint32_t cout_valid) {
    __shared__ v4s ws[27 * 16 * 4];   // [tap][cout][q]: cin 4q .. 4q + 3
    const v4s* Wv = reinterpret_cast<const v4s*>(W);
    for (uint32_t i = threadIdx.x; i < 27u * 16u * 4u; i += 256u) ws[i] = Wv[i];
    __syncthreads();
    const uint32_t nzt = (n2 + 15u) / 16u;
    const uint32_t wid = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (wid >= n0 * n1 * nzt) return;
    const uint32_t zt = wid % nzt, y = (wid / nzt) % n1, x = wid / (nzt * n1);
    const uint32_t l = threadIdx.x & 63u, q = l >> 4, r = l & 15u;
    const uint32_t z0 = zt * 16u;
    v4s a[27];
#pragma unroll
    for (int t = 0; t < 27; ++t) {
        const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
        const int xs = (int)x + dx, ys = (int)y + dy, zs = (int)(z0 + r) + dz;
        a[t] = v4s{0, 0, 0, 0};
        if (xs >= 0 && xs < (int)n0 && ys >= 0 && ys < (int)n1 && zs >= 0 && zs < (int)n2)
            a[t] = *reinterpret_cast<const v4s*>(X + (((size_t)xs * n1 + ys) * n2 + zs) * 16u + 4u * q);
    }
    v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int t = 0; t < 27; ++t)
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[t], ws[(t * 16 + (int)r) * 4 + (int)q], acc, 0, 0, 0);
    const uint32_t cout = r;
    const bool valid = cout < cout_valid;
    const float bv = (bias && valid) ? bias[cout] : 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t z = z0 + 4u * q + (uint32_t)i;
        if (z >= n2) continue;
        float v = valid ? acc[i] + bv : 0.0f;  // channels >= cout_valid are written as 0
        if (act == 1) v = v > 0.0f ? v : 0.0f;
        const size_t o = (((size_t)x * n1 + y) * n2 + z) * 16u + cout;
        if (yb) yb[o] = to_bf16(v);
        if (yf) yf[o] = v;
    }
}

hipError_t launch_conv3_fwd(const uint16_t* X, const uint16_t* W, const float* bias, uint16_t* yb,
                            float* yf, uint32_t n0, uint32_t n1, uint32_t n2, int act,
                            uint32_t cout_valid, hipStream_t s) {
    const uint32_t waves = n0 * n1 * ((n2 + 15u) / 16u);
    hipLaunchKernelGGL(conv3_fwd_kernel, dim3((waves + 3u) / 4u), dim3(256), 0, s, X, W, bias, yb, yf, n0,
                       n1, n2, act, cout_valid);
    return hipGetLastError();
}

// ------------------------------------------------------------------ weight gradient
// dW[t][cout][cin] += sum_v dZ[v][cout] * X[v + off(t)][cin]; each wave takes voxel chunks of
// 16 (the MFMA K dimension) with a grid stride and accumulates all 27 taps in registers.
__global__ __launch_bounds__(256) void conv3_dw_kernel(const uint16_t* __restrict__ dZ,
                                                       const uint16_t* __restrict__ X,
                                                       float* __restrict__ dW, float* __restrict__ db,
                                                       uint32_t n0, uint32_t n1, uint32_t n2) {
    __shared__ float sdw[27 * 256];  // the work-group's partial dW (one global atomic per element)
    __shared__ float sdb[16];
    for (uint32_t i = threadIdx.x; i < 27u * 256u; i += blockDim.x) sdw[i] = 0.0f;
    if (threadIdx.x < 16u) sdb[threadIdx.x] = 0.0f;
    __syncthreads();
    const uint32_t V = n0 * n1 * n2;
    const uint32_t nchunks = (V + 15u) / 16u;
    const uint32_t wid = blockIdx.x * 4u + (threadIdx.x >> 6), nw = gridDim.x * 4u;
    const uint32_t l = threadIdx.x & 63u, q = l >> 4, r = l & 15u;
    v4f acc[27];
#pragma unroll
    for (int t = 0; t < 27; ++t) acc[t] = (v4f){0.0f, 0.0f, 0.0f, 0.0f};
    float bsum = 0.0f;
    for (uint32_t ch = wid; ch < nchunks; ch += nw) {
        int vx[4], vy[4], vz[4];
        bool ok[4];
        v4s a;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t v = ch * 16u + 4u * q + (uint32_t)i;
            ok[i] = v < V;
            const uint32_t vv = ok[i] ? v : 0u;
            vz[i] = (int)(vv % n2);
            vy[i] = (int)((vv / n2) % n1);
            vx[i] = (int)(vv / (n2 * n1));
            const short dz = ok[i] ? (short)dZ[(size_t)vv * 16u + r] : (short)0;
            a[i] = dz;
            bsum += ok[i] ? from_bf16((uint16_t)dz) : 0.0f;
        }
#pragma unroll
        for (int t = 0; t < 27; ++t) {
            const int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dzz = t % 3 - 1;
            v4s b;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int xs = vx[i] + dx, ys = vy[i] + dy, zs = vz[i] + dzz;
                short val = 0;
                if (ok[i] && xs >= 0 && xs < (int)n0 && ys >= 0 && ys < (int)n1 && zs >= 0 && zs < (int)n2)
                    val = (short)X[(((size_t)xs * n1 + ys) * n2 + zs) * 16u + r];
                b[i] = val;
            }
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
        }
    }
    // D lane layout: rows (cout) 4q + i, column (cin) r; waves of the group add in LDS
#pragma unroll
    for (int t = 0; t < 27; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i)
            atomicAdd(&sdw[((uint32_t)t * 16u + 4u * q + (uint32_t)i) * 16u + r], acc[t][i]);
    // bias: lane (q, r) summed dZ[.][r] over its voxels
    for (int off = 16; off < 64; off <<= 1) bsum += __shfl_xor(bsum, off, 64);
    if (q == 0) atomicAdd(&sdb[r], bsum);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < 27u * 256u; i += blockDim.x) atomicAdd(&dW[i], sdw[i]);
    if (threadIdx.x < 16u && db) atomicAdd(&db[threadIdx.x], sdb[threadIdx.x]);
}

hipError_t launch_conv3_dw(const uint16_t* dZ, const uint16_t* X, float* dW, float* db, uint32_t n0,
                           uint32_t n1, uint32_t n2, hipStream_t s) {
    const uint32_t V = n0 * n1 * n2;
    const uint32_t nchunks = (V + 15u) / 16u;
    uint32_t waves = (nchunks + 63u) / 64u;  // ~64 chunks (1024 voxels) per wave
    if (waves < 1u) waves = 1u;
    hipLaunchKernelGGL(conv3_dw_kernel, dim3((waves + 3u) / 4u), dim3(256), 0, s, dZ, X, dW, db, n0, n1, n2);
    return hipGetLastError();
}

// ------------------------------------------------------------------ elementwise
// Flipped/transposed weights for the data gradient: Wt[t'][cin][cout] = W[26 - t'][cout][cin].
__global__ void flip_transpose_kernel(const uint16_t* __restrict__ W, uint16_t* __restrict__ Wt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 27u * 256u) return;
    const uint32_t t = i / 256u, a = (i / 16u) % 16u, b = i % 16u;  // Wt[t][a = cin][b = cout]
    Wt[i] = W[((26u - t) * 16u + b) * 16u + a];
}

hipError_t launch_flip_transpose(const uint16_t* W, uint16_t* Wt, hipStream_t s) {
    hipLaunchKernelGGL(flip_transpose_kernel, dim3((27u * 256u + 255u) / 256u), dim3(256), 0, s, W, Wt);
    return hipGetLastError();
}

// dZ_prev = dX * (Y_prev > 0) converted to bf16 (ReLU backward into the previous layer).
__global__ void relu_back_kernel(const float* __restrict__ dX, const uint16_t* __restrict__ Y,
                                 uint16_t* __restrict__ dZ, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dZ[i] = from_bf16(Y[i]) > 0.0f ? to_bf16(dX[i]) : (uint16_t)0;
}

hipError_t launch_relu_back(const float* dX, const uint16_t* Y, uint16_t* dZ, uint64_t n, hipStream_t s) {
    hipLaunchKernelGGL(relu_back_kernel, dim3((uint32_t)((n + 255u) / 256u)), dim3(256), 0, s, dX, Y, dZ, n);
    return hipGetLastError();
}

// Weighted binary cross-entropy on logits [V][16] (channels < nout) against labels [V][nout]
// (0/1 bytes), only where mask[v] != 0.  dlogits in bf16 (zero elsewhere), loss sum in double.
__global__ __launch_bounds__(256) void bce_kernel(const float* __restrict__ logits, const uint8_t* __restrict__ labels,
                           const uint8_t* __restrict__ mask, uint32_t V, uint32_t nout, float pos_weight,
                           float scale, uint16_t* __restrict__ dlogits, double* __restrict__ loss) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0.0;
    if (i < V * 16u) {
        const uint32_t v = i / 16u, c = i % 16u;
        float g = 0.0f;
        if (c < nout && (!mask || mask[v])) {
            const float z = logits[i];
            const float y = (float)labels[(size_t)v * nout + c];
            const float sg = 1.0f / (1.0f + __expf(-z));
            // log(sigmoid(z)) and log(1 - sigmoid(z)) in a stable form
            const float lsp = z >= 0.0f ? -log1pf(__expf(-z)) : z - log1pf(__expf(z));
            const float lsn = lsp - z;
            l = -(double)(pos_weight * y * lsp + (1.0f - y) * lsn) * scale;
            g = (sg * (pos_weight * y + 1.0f - y) - pos_weight * y) * scale;
        }
        dlogits[i] = to_bf16(g);
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off, 64);
    // one atomic per work-group (a double atomic per wave contends on one address)
    __shared__ double wsum[4];
    if ((threadIdx.x & 63u) == 0u) wsum[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double t = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
        if (t != 0.0) atomicAdd(loss, t);
    }
}

hipError_t launch_bce(const float* logits, const uint8_t* labels, const uint8_t* mask, uint32_t V,
                      uint32_t nout, float pos_weight, float scale, uint16_t* dlogits, double* loss,
                      hipStream_t s) {
    const uint32_t n = V * 16u;
    hipLaunchKernelGGL(bce_kernel, dim3((n + 255u) / 256u), dim3(256), 0, s, logits, labels, mask, V, nout,
                       pos_weight, scale, dlogits, loss);
    return hipGetLastError();
}

// Adam on fp32 master parameters; refreshes the bf16 copies of the weights.
__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, uint16_t* __restrict__ pb, uint32_t n, uint32_t nb,
                            float lr, float b1, float b2, float eps, float c1, float c2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float upd = lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] -= upd;
    if (i < nb) pb[i] = to_bf16(p[i]);
}

hipError_t launch_adam(float* p, const float* g, float* m, float* v, uint16_t* pb, uint32_t n, uint32_t nb,
                       float lr, float b1, float b2, float eps, float c1, float c2, hipStream_t s) {
    hipLaunchKernelGGL(adam_kernel, dim3((n + 255u) / 256u), dim3(256), 0, s, p, g, m, v, pb, n, nb, lr, b1, b2,
                       eps, c1, c2);
    return hipGetLastError();
}

__global__ void to_bf16_kernel(const float* __restrict__ src, uint16_t* __restrict__ dst, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = to_bf16(src[i]);
}

hipError_t launch_to_bf16(const float* src, uint16_t* dst, uint64_t n, hipStream_t s) {
    hipLaunchKernelGGL(to_bf16_kernel, dim3((uint32_t)((n + 255u) / 256u)), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

// ------------------------------------------------------------------ octree features / labels
// Dense level grid of input features (bf16 [V][16]) from the level's evaluated block list
// (grid cells never evaluated stay 0):
//   c0 = U / running max, c1 = V / running max, c2 = 1 (evaluated), c3 = U / max U,
//   c4 = level_code (log2 of the block edge / 8); others 0.  mask[v] = 1 where evaluated.
__global__ void oct_features_kernel(const uint32_t* __restrict__ blocks, const float* __restrict__ U,
                                    const float* __restrict__ Vv, uint32_t n, const double* __restrict__ stats,
                                    uint32_t n1, uint32_t n2, float level_code, uint16_t* __restrict__ X,
                                    uint8_t* __restrict__ mask) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = blocks[i];
    const uint32_t x = c & 1023u, y = (c >> 10) & 1023u, z = c >> 20;
    const size_t v = ((size_t)x * n1 + y) * n2 + z;
    const size_t o = v * 16u;
    const double mrun = stats[0] > 0.0 ? stats[0] : 1.0;
    const double umax = stats[1] > 0.0 ? stats[1] : 1.0;
    X[o + 0] = to_bf16((float)((double)U[i] / mrun));
    X[o + 1] = to_bf16((float)((double)Vv[i] / mrun));
    X[o + 2] = to_bf16(1.0f);
    X[o + 3] = to_bf16((float)((double)U[i] / umax));
    X[o + 4] = to_bf16(level_code);
    if (mask) mask[v] = 1u;
}

hipError_t launch_oct_features(const uint32_t* blocks, const float* U, const float* V, uint32_t n,
                               const double* stats, uint32_t n1, uint32_t n2, float level_code, uint16_t* X,
                               uint8_t* mask, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(oct_features_kernel, dim3((n + 255u) / 256u), dim3(256), 0, s, blocks, U, V, n, stats, n1,
                       n2, level_code, X, mask);
    return hipGetLastError();
}

// 2x2x2 max pooling: out (m0, m1, m2) from in (2m0, 2m1, 2m2).
__global__ void maxpool2_kernel(const float* __restrict__ in, float* __restrict__ out, uint32_t m0,
                                uint32_t m1, uint32_t m2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m0 * m1 * m2) return;
    const uint32_t z = i % m2, y = (i / m2) % m1, x = i / (m2 * m1);
    const uint32_t N1 = 2u * m1, N2 = 2u * m2;
    float m = -INFINITY;
    for (uint32_t c = 0; c < 8u; ++c) {
        const uint32_t xi = 2u * x + ((c >> 2) & 1u), yi = 2u * y + ((c >> 1) & 1u), zi = 2u * z + (c & 1u);
        m = fmaxf(m, in[((size_t)xi * N1 + yi) * N2 + zi]);
    }
    out[i] = m;
}

hipError_t launch_maxpool2(const float* in, float* out, uint32_t m0, uint32_t m1, uint32_t m2, hipStream_t s) {
    const uint32_t n = m0 * m1 * m2;
    hipLaunchKernelGGL(maxpool2_kernel, dim3((n + 255u) / 256u), dim3(256), 0, s, in, out, m0, m1, m2);
    return hipGetLastError();
}

// labels[v][c] (level grid n, 8 children) = pooled(child) >= thr; pooled is the max-pooled
// fine volume at the children's resolution (2n0, 2n1, 2n2).  thr = frac * (*vmax).
__global__ void child_labels_kernel(const float* __restrict__ pooled, uint32_t n0, uint32_t n1, uint32_t n2,
                                    const double* __restrict__ vmax, float frac, uint8_t* __restrict__ labels) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n0 * n1 * n2 * 8u) return;
    const uint32_t v = i / 8u, c = i % 8u;
    const uint32_t z = v % n2, y = (v / n2) % n1, x = v / (n2 * n1);
    const uint32_t xi = 2u * x + ((c >> 2) & 1u), yi = 2u * y + ((c >> 1) & 1u), zi = 2u * z + (c & 1u);
    const float p = pooled[((size_t)xi * (2u * n1) + yi) * (2u * n2) + zi];
    labels[i] = (double)p >= (double)frac * vmax[0] ? 1u : 0u;
}

hipError_t launch_child_labels(const float* pooled, uint32_t n0, uint32_t n1, uint32_t n2, const double* vmax,
                               float frac, uint8_t* labels, hipStream_t s) {
    const uint32_t n = n0 * n1 * n2 * 8u;
    hipLaunchKernelGGL(child_labels_kernel, dim3((n + 255u) / 256u), dim3(256), 0, s, pooled, n0, n1, n2, vmax,
                       frac, labels);
    return hipGetLastError();
}

// FP64 max of a float array in two passes: every work-group writes the max of its grid-stride
// share to partial[blockIdx], then one work-group reduces the partials into out[0].
__global__ __launch_bounds__(256) void max_f64_kernel(const float* __restrict__ x, uint64_t n,
                                                      double* __restrict__ out) {
    __shared__ double sh[256];
    double m = -INFINITY;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        m = fmax(m, (double)x[i]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (uint32_t o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void max_f64_finish_kernel(const double* __restrict__ partial, uint32_t n,
                                                             double* __restrict__ out) {
    __shared__ double sh[256];
    double m = -INFINITY;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) m = fmax(m, partial[i]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (uint32_t o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

hipError_t launch_max_f64(const float* x, uint64_t n, double* partial, uint32_t max_parts, double* out,
                          hipStream_t s) {
    uint32_t parts = (uint32_t)std::min<uint64_t>((n + 255u) / 256u, max_parts);
    if (parts == 0u) parts = 1u;
    hipLaunchKernelGGL(max_f64_kernel, dim3(parts), dim3(256), 0, s, x, n, partial);
    hipLaunchKernelGGL(max_f64_finish_kernel, dim3(1), dim3(256), 0, s, partial, parts, out);
    return hipGetLastError();
}

// Learned child flags of a level: flag[8i + c] = (U[i] >= threshold) && sigmoid(logit) > p,
// logits read from the dense level grid at the block's position.
// Children of block b: 1 = refined, 2 = computed at full resolution (every fine voxel), 0 =
// dropped.  Without a net: 1 iff U >= tau x running max (stats[4]).  With a net, a child that
// passes that rule is refined iff its logit exceeds logit_thr or U >= keep x running max
// (stats[0]; keep 0 never overrides); below logit_thr but above fallback_thr (a low-confidence
// child) it takes the full-resolution fallback (fallback_thr >= logit_thr: never).
__global__ void oct_child_flags_kernel(const uint32_t* __restrict__ blocks, const float* __restrict__ U,
                                       uint32_t n, const double* __restrict__ stats,
                                       const float* __restrict__ logits, uint32_t n1, uint32_t n2,
                                       float logit_thr, float fallback_thr, float keep,
                                       uint32_t* __restrict__ flags) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * 8u) return;
    const uint32_t b = i / 8u, c = i % 8u;
    uint32_t f = (double)U[b] >= stats[4] ? 1u : 0u;
    if (logits && f) {
        const uint32_t cc = blocks[b];
        const uint32_t x = cc & 1023u, y = (cc >> 10) & 1023u, z = cc >> 20;
        const float lg = logits[(((size_t)x * n1 + y) * n2 + z) * 16u + c];
        const bool keep_bound = keep > 0.0f && (double)U[b] >= (double)keep * stats[0];
        if (!(lg > logit_thr) && !keep_bound) f = lg > fallback_thr ? 2u : 0u;
    }
    flags[i] = f;
}

hipError_t launch_oct_child_flags(const uint32_t* blocks, const float* U, uint32_t n, const double* stats,
                                  const float* logits, uint32_t n1, uint32_t n2, float logit_thr, float fallback_thr,
                                  float keep, uint32_t* flags, hipStream_t s) {
    const uint32_t m = n * 8u;
    if (m == 0) return hipSuccess;
    hipLaunchKernelGGL(oct_child_flags_kernel, dim3((m + 255u) / 256u), dim3(256), 0, s, blocks, U, n, stats,
                       logits, n1, n2, logit_thr, fallback_thr, keep, flags);
    return hipGetLastError();
}

}  // namespace bpx
