// Octree coarse-to-fine backprojection (confocal captures, fp32 samples).
//
// Level l of L refinements holds blocks of B = 2^(L-l) fine voxels per axis.  For a block the
// kernels compute
//   U = sum_p max{ H[p][b] : b in the block's bin window for pair p }   (>= every fine voxel's
//       value inside the block when samples are non-negative: a sound upper bound), and
//   V = the exact single-tap value of one fine voxel of the block (a lower bound of the max).
// The bin window comes from the block's bounding sphere (centre m, radius R): every fine voxel
// v in the block has |w - v| in [|w - m| - R, |w - m| + R], so its bin lies in a window of
// fixed width W_l starting at floor(2(|w - m| - R)/dT + c) - margin.  A per-level sliding
// window maximum table M_l[p][b] = max(H[p][b .. b + W_l - 1], 0) turns the window max into one
// read.  Running max and L2 norm of each level are reduced in FP64; blocks with
// U >= tau * max_so_far are refined.  The finest level is computed exactly (same fused
// sample+threshold layout and the same per-update arithmetic as the dense kernel).
#include <algorithm>

#include "bpx_device.hpp"
#include "bpx_internal.hpp"

namespace bpx {

// ------------------------------------------------------------------ sliding-window max
// out[p][b] = max(0, max(in[p][b .. b + w - 1])) for b < Tc, out[p][Tc] = 0; samples beyond
// the row count as 0 (an out-of-range bin contributes nothing).  in_pitch / out_pitch in
// floats.  Built by doubling: width 2w from two width-w tables.
__global__ __launch_bounds__(256) void winmax_double_kernel(const float* __restrict__ in, uint32_t in_pitch,
                                                            float* __restrict__ out, uint32_t out_pitch,
                                                            uint32_t rows, uint32_t Tc, uint32_t half,
                                                            const uint32_t* __restrict__ row_index) {
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = Tc + 1u;
    if (gid >= (uint64_t)rows * n) return;
    const uint32_t p = (uint32_t)(gid / n), b = (uint32_t)(gid % n);
    const float* r = in + (size_t)(row_index ? row_index[p] : p) * in_pitch;
    float m = 0.0f;
    if (b < Tc) {
        m = fmaxf(m, r[b]);
        if (b + half < Tc) m = fmaxf(m, r[b + half]);
    }
    out[(size_t)p * out_pitch + b] = m;
}

// out[b] = max(t2[b], t2[b + w - p2]) where t2 is the width-p2 table (p2 <= w < 2 p2).
__global__ __launch_bounds__(256) void winmax_combine_kernel(const float* __restrict__ t2, uint32_t pitch2,
                                                             float* __restrict__ out, uint32_t out_pitch,
                                                             uint32_t rows, uint32_t Tc, uint32_t shift) {
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = Tc + 1u;
    if (gid >= (uint64_t)rows * n) return;
    const uint32_t p = (uint32_t)(gid / n), b = (uint32_t)(gid % n);
    const float* r = t2 + (size_t)p * pitch2;
    float m = r[b];
    const uint32_t b2 = b + shift;
    if (b2 < Tc) m = fmaxf(m, r[b2]);
    out[(size_t)p * out_pitch + b] = m;
}

// One work-group per row: the width-1 table max(0, H) goes to LDS, widths double in LDS
// (ping-pong), and each level's table is written out when its width is reached, with the same
// arithmetic as winmax_double_kernel / winmax_combine_kernel.
constexpr uint32_t kWinmaxLdsBytes = 64u * 1024u;

bool winmax_fused_fits(uint32_t Tc) { return 2ull * 4ull * ((uint64_t)Tc + 1u) <= kWinmaxLdsBytes; }

// rows: the source row of table row p (nullptr = row p).
__global__ __launch_bounds__(1024) void winmax_fused_kernel(const float* __restrict__ in, uint32_t in_pitch,
                                                            uint32_t Tc, uint32_t out_pitch, WinmaxLevels lv,
                                                            const uint32_t* __restrict__ rows) {
    extern __shared__ float lds[];
    float* a = lds;
    float* b = lds + (Tc + 1u);
    const uint32_t p = blockIdx.x;
    const float* r = in + (size_t)(rows ? rows[p] : p) * in_pitch;
    const uint32_t n = Tc + 1u;
    for (uint32_t t = threadIdx.x; t < n; t += blockDim.x) a[t] = t < Tc ? fmaxf(0.0f, r[t]) : 0.0f;
    __syncthreads();
    uint32_t wmax = 1u;
    for (uint32_t k = 0; k < lv.count; ++k) wmax = max(wmax, lv.W[k]);
    for (uint32_t w = 1u;; w *= 2u) {
        for (uint32_t k = 0; k < lv.count; ++k) {
            if (lv.W[k] < w || lv.W[k] >= 2u * w) continue;
            const uint32_t shift = lv.W[k] - w;
            float* o = lv.out[k] + (size_t)p * out_pitch;
            for (uint32_t t = threadIdx.x; t < n; t += blockDim.x) {
                float m = a[t];
                if (t + shift < Tc) m = fmaxf(m, a[t + shift]);
                o[t] = m;
            }
        }
        if (2u * w > wmax) break;
        for (uint32_t t = threadIdx.x; t < n; t += blockDim.x) {
            float m = 0.0f;
            if (t < Tc) {
                m = fmaxf(m, a[t]);
                if (t + w < Tc) m = fmaxf(m, a[t + w]);
            }
            b[t] = m;
        }
        __syncthreads();
        float* tmp = a;
        a = b;
        b = tmp;
    }
}

hipError_t launch_winmax_fused(const float* in, uint32_t in_pitch, uint32_t rows, uint32_t Tc, uint32_t out_pitch,
                               const WinmaxLevels& lv, const uint32_t* row_index, hipStream_t s) {
    if (!winmax_fused_fits(Tc) || lv.count == 0 || lv.count > 8) return hipErrorInvalidValue;
    if (rows == 0) return hipSuccess;
    const uint32_t lds = 2u * 4u * (Tc + 1u);
    hipLaunchKernelGGL(winmax_fused_kernel, dim3(rows), dim3(1024), lds, s, in, in_pitch, Tc, out_pitch, lv,
                       row_index);
    return hipGetLastError();
}

hipError_t launch_winmax_double(const float* in, uint32_t in_pitch, float* out, uint32_t out_pitch,
                                uint32_t rows, uint32_t Tc, uint32_t half, const uint32_t* row_index, hipStream_t s) {
    const uint64_t n = (uint64_t)rows * (Tc + 1u);
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(winmax_double_kernel, dim3((uint32_t)((n + 255u) / 256u)), dim3(256), 0, s, in,
                       in_pitch, out, out_pitch, rows, Tc, half, row_index);
    return hipGetLastError();
}

hipError_t launch_winmax_combine(const float* t2, uint32_t pitch2, float* out, uint32_t out_pitch,
                                 uint32_t rows, uint32_t Tc, uint32_t shift, hipStream_t s) {
    const uint64_t n = (uint64_t)rows * (Tc + 1u);
    hipLaunchKernelGGL(winmax_combine_kernel, dim3((uint32_t)((n + 255u) / 256u)), dim3(256), 0, s, t2,
                       pitch2, out, out_pitch, rows, Tc, shift);
    return hipGetLastError();
}

// ------------------------------------------------------------------ level bounds
__device__ __forceinline__ void unpack_block(uint32_t c, uint32_t& x, uint32_t& y, uint32_t& z) {
    x = c & 1023u;
    y = (c >> 10) & 1023u;
    z = c >> 20;
}

// K blocks per lane (consecutive list entries: siblings, so their bins share cache lines);
// pairs in the wave-uniform loop (pair data in SGPRs).  Coarse levels have few blocks, so the
// pair range is split into segments (blockIdx.y) whose partial sums are added in segment
// order by oct_bound_combine_kernel (deterministic).
// FLAT: every pair's wall point has z == a.lz, so the z terms of the squared distances are
// per-block constants ((lz - z)^2, the same rounded value the full expression forms).
template <int K, bool FLAT>
__global__ __launch_bounds__(256) void oct_bound_kernel(OctBoundArgs a) {
    const uint32_t base = (blockIdx.x * blockDim.x + threadIdx.x) * K;
    const double h = 0.5 * (double)(a.B - 1u);
    const uint32_t rep = a.B / 2u;
    float mx[K], my[K], mz[K], vx[K], vy[K], vz[K];
#pragma unroll
    for (int r = 0; r < K; ++r) {
        uint32_t bx, by, bz;
        unpack_block(a.blocks[min(base + r, a.nblocks - 1u)], bx, by, bz);
        // block centre (double precision, then float) and representative fine voxel
        float m[3], v[3];
        for (int k = 0; k < 3; ++k)
            m[k] = (float)((double)a.g.zero[k] + a.g.inc[k] * ((double)bx * a.B + h) +
                           a.g.inc[3 + k] * ((double)by * a.B + h) + a.g.inc[6 + k] * ((double)bz * a.B + h));
        const float fx = (float)(bx * a.B + rep), fy = (float)(by * a.B + rep), fz = (float)(bz * a.B + rep);
        for (int k = 0; k < 3; ++k) {
            float t = a.g.zero[k] + a.g.inc[k] * fx;
            t = t + a.g.inc[3 + k] * fy;
            v[k] = t + a.g.inc[6 + k] * fz;
        }
        mx[r] = m[0];
        my[r] = m[1];
        mz[r] = m[2];
        vx[r] = v[0];
        vy[r] = v[1];
        vz[r] = v[2];
    }
    float mz2[K], vz2[K];
#pragma unroll
    for (int r = 0; r < K; ++r) {
        const float cm = a.lz - mz[r], cv = a.lz - vz[r];
        mz2[r] = cm * cm;
        vz2[r] = cv * cv;
    }
    const auto mr = make_rsrc(a.M, a.m_bytes);
    const auto er = make_rsrc(a.E, a.e_bytes);
    const float fTc = (float)a.Tc;
    using Off = ElemOffset<8>;
    const float top = Off::kBias + fTc;
    float U[K], Vs[K];
#pragma unroll
    for (int r = 0; r < K; ++r) U[r] = Vs[r] = 0.0f;
    const uint32_t p0 = blockIdx.y * a.chunk, p1 = min(a.P, p0 + a.chunk);
    for (uint32_t p = p0; p < p1; ++p) {
        const float4 pr = a.pairs[p];
        // table row p of the level's pair set; fused sample row of the same pair in the capture
        const uint32_t moff = p * a.ldm * 4u, eoff = (a.erow ? a.erow[p] : p) * a.lde * 8u;
        const float c1R = pr.w + a.cR;
        const float wb = pr.w + Off::kBias;
#pragma unroll
        for (int r = 0; r < K; ++r) {
            // window start for the bounding sphere
            float sm, s;
            if (FLAT) {
                const float mx_ = pr.x - mx[r], my_ = pr.y - my[r], vx_ = pr.x - vx[r], vy_ = pr.y - vy[r];
                sm = (mx_ * mx_ + my_ * my_) + mz2[r];
                s = (vx_ * vx_ + vy_ * vy_) + vz2[r];
            } else {
                sm = sq3(pr.x - mx[r], pr.y - my[r], pr.z - mz[r]);
                s = sq3(pr.x - vx[r], pr.y - vy[r], pr.z - vz[r]);
            }
            const float qlo = __builtin_fmaf(__builtin_amdgcn_sqrtf(sm), a.inv2, c1R);
            const uint32_t blo = (uint32_t)(int)__builtin_amdgcn_fmed3f(qlo, 0.0f, fTc);
            U[r] = U[r] + buf_f32(mr, blo * 4u, moff);
            // exact tap of the representative fine voxel (element offset from the biased
            // quotient: within 1/8 of q1 = fma(sqrt(s), inv2, w), the texture path's budget)
            const float y = __builtin_fmaf(__builtin_amdgcn_sqrtf(s), a.inv2, wb);
            Vs[r] = Vs[r] + FusedF32::pick(er, Off::clamped(y, top), eoff, s);
        }
    }
#pragma unroll
    for (int r = 0; r < K; ++r) {
        const uint32_t i = base + r;
        if (i >= a.nblocks) continue;
        if (a.partU) {
            a.partU[(size_t)blockIdx.y * a.nblocks + i] = U[r];
            a.partV[(size_t)blockIdx.y * a.nblocks + i] = Vs[r];
        } else {
            a.U[i] = U[r];
            a.V[i] = Vs[r];
        }
    }
}

__global__ __launch_bounds__(256) void oct_bound_combine_kernel(const float* __restrict__ partU,
                                                                const float* __restrict__ partV, uint32_t n,
                                                                uint32_t segments, float* __restrict__ U,
                                                                float* __restrict__ V) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float u = 0.0f, v = 0.0f;
    for (uint32_t s = 0; s < segments; ++s) {
        u = u + partU[(size_t)s * n + i];
        v = v + partV[(size_t)s * n + i];
    }
    U[i] = u;
    V[i] = v;
}

// Work-groups of 256 lanes of `kernel` resident at once on the device (occupancy x CUs).
template <class F>
static uint32_t resident_groups(F kernel, uint32_t& cache) {
    if (!cache) {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) != hipSuccess) {
            (void)hipGetLastError();
            return 2048u;
        }
        cache = (uint32_t)std::max(1, cus * per_cu);
    }
    return cache;
}

// Each lane loops over all its pairs, so a launch takes ~ (rounds of resident work-groups) x
// (items per lane); a last round that is only partly filled costs a whole round.  Returns the
// items per lane among `ks` (ascending) minimising rounds x items; ties go to the smaller,
// whose extra lanes hide more gather latency (measured: one block per lane instead of two
// for 534K level-3 blocks x 4,096 pairs took the whole learned 256^3 octree 16.4 -> 16.0 ms).
static int pick_items_per_lane(uint64_t n, const int* ks, const uint32_t* slots, int count) {
    int best = ks[0];
    uint64_t best_cost = ~0ull;
    for (int i = 0; i < count; ++i) {
        const uint64_t wgs = (n + 256u * (uint64_t)ks[i] - 1) / (256u * (uint64_t)ks[i]);
        const uint64_t cost = (wgs + slots[i] - 1) / slots[i] * (uint64_t)ks[i];
        if (cost < best_cost) {
            best_cost = cost;
            best = ks[i];
        }
    }
    return best;
}

uint32_t oct_bound_segments(uint32_t nblocks, uint32_t P, uint64_t part_capacity) {
    // aim for >= 2^18 lanes in flight, segments of >= 64 pairs, partials within capacity
    if (nblocks == 0 || P == 0) return 1u;
    uint64_t s = ((1ull << 18) + nblocks - 1) / nblocks;
    s = std::min<uint64_t>(s, std::max<uint32_t>(P / 64u, 1u));
    s = std::min<uint64_t>(s, part_capacity / nblocks);
    return (uint32_t)std::max<uint64_t>(s, 1u);
}

hipError_t launch_oct_bound(const OctBoundArgs& a_in, uint32_t segments, hipStream_t s) {
    if (a_in.nblocks == 0) return hipSuccess;
    OctBoundArgs a = a_in;
    if (segments <= 1u || !a.partU || !a.partV) {
        a.chunk = a.P;
        a.partU = a.partV = nullptr;
        segments = 1u;
    } else {
        a.chunk = (a.P + segments - 1u) / segments;
        segments = (a.P + a.chunk - 1u) / a.chunk;
    }
    const bool flat = a.flat != 0u;
    if (segments > 1u) {
        const dim3 grid((a.nblocks + 255u) / 256u, segments);
        if (flat) hipLaunchKernelGGL((oct_bound_kernel<1, true>), grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((oct_bound_kernel<1, false>), grid, dim3(256), 0, s, a);
    } else {
        static uint32_t c1 = 0, c2 = 0;
        const int ks[2] = {1, 2};
        const uint32_t slots[2] = {resident_groups(oct_bound_kernel<1, true>, c1),
                                   resident_groups(oct_bound_kernel<2, true>, c2)};
        const int K = pick_items_per_lane(a.nblocks, ks, slots, 2);
        const uint32_t lanes = (a.nblocks + K - 1u) / K;
        const dim3 grid((lanes + 255u) / 256u, 1u);
        if (K == 1) {
            if (flat) hipLaunchKernelGGL((oct_bound_kernel<1, true>), grid, dim3(256), 0, s, a);
            else hipLaunchKernelGGL((oct_bound_kernel<1, false>), grid, dim3(256), 0, s, a);
        } else {
            if (flat) hipLaunchKernelGGL((oct_bound_kernel<2, true>), grid, dim3(256), 0, s, a);
            else hipLaunchKernelGGL((oct_bound_kernel<2, false>), grid, dim3(256), 0, s, a);
        }
    }
    if (segments > 1u)
        hipLaunchKernelGGL(oct_bound_combine_kernel, dim3((a.nblocks + 255u) / 256u), dim3(256), 0, s, a.partU,
                           a.partV, a.nblocks, segments, a.U, a.V);
    return hipGetLastError();
}

// ------------------------------------------------------------------ FP64 reductions
// stats[0] = running max of V (kept across levels), stats[1] = max of U, stats[2] = sum U^2,
// stats[3] = sum V^2, stats[4] = threshold tau * stats[0] (as used by the selection).
__global__ __launch_bounds__(256) void oct_reduce_kernel(const float* __restrict__ U, const float* __restrict__ V,
                                                         uint32_t n, double* __restrict__ partial) {
    __shared__ double sh[4][256];
    double mv = -INFINITY, mu = -INFINITY, su = 0.0, sv = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double u = U[i], v = V[i];
        mv = fmax(mv, v);
        mu = fmax(mu, u);
        su += u * u;
        sv += v * v;
    }
    sh[0][threadIdx.x] = mv;
    sh[1][threadIdx.x] = mu;
    sh[2][threadIdx.x] = su;
    sh[3][threadIdx.x] = sv;
    __syncthreads();
    for (uint32_t o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            sh[0][threadIdx.x] = fmax(sh[0][threadIdx.x], sh[0][threadIdx.x + o]);
            sh[1][threadIdx.x] = fmax(sh[1][threadIdx.x], sh[1][threadIdx.x + o]);
            sh[2][threadIdx.x] += sh[2][threadIdx.x + o];
            sh[3][threadIdx.x] += sh[3][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) partial[4 * blockIdx.x + k] = sh[k][0];
}

__global__ __launch_bounds__(256) void oct_finish_kernel(const double* __restrict__ partial, uint32_t nparts,
                                                         double* __restrict__ stats, double tau) {
    __shared__ double sh[4][256];
    double mv = -INFINITY, mu = -INFINITY, su = 0.0, sv = 0.0;
    for (uint32_t i = threadIdx.x; i < nparts; i += blockDim.x) {
        mv = fmax(mv, partial[4 * i]);
        mu = fmax(mu, partial[4 * i + 1]);
        su += partial[4 * i + 2];
        sv += partial[4 * i + 3];
    }
    sh[0][threadIdx.x] = mv;
    sh[1][threadIdx.x] = mu;
    sh[2][threadIdx.x] = su;
    sh[3][threadIdx.x] = sv;
    __syncthreads();
    for (uint32_t o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            sh[0][threadIdx.x] = fmax(sh[0][threadIdx.x], sh[0][threadIdx.x + o]);
            sh[1][threadIdx.x] = fmax(sh[1][threadIdx.x], sh[1][threadIdx.x + o]);
            sh[2][threadIdx.x] += sh[2][threadIdx.x + o];
            sh[3][threadIdx.x] += sh[3][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        stats[0] = fmax(stats[0], sh[0][0]);
        stats[1] = sh[1][0];
        stats[2] = sh[2][0];
        stats[3] = sh[3][0];
        stats[4] = tau * stats[0];
    }
}

hipError_t launch_oct_reduce(const float* U, const float* V, uint32_t n, double* partial,
                             uint32_t max_parts, double* stats, double tau, hipStream_t s) {
    uint32_t parts = (n + 255u) / 256u;
    if (parts > max_parts) parts = max_parts;
    if (parts == 0) parts = 1;
    hipLaunchKernelGGL(oct_reduce_kernel, dim3(parts), dim3(256), 0, s, U, V, n, partial);
    hipLaunchKernelGGL(oct_finish_kernel, dim3(1), dim3(256), 0, s, partial, parts, stats, tau);
    return hipGetLastError();
}

// ------------------------------------------------------------------ selection
// Exclusive scan of flags in chunks of 1024 (one work-group each), chunk totals in sums.
// Items i with flags[i] == want are counted.
__global__ __launch_bounds__(1024) void scan_chunk_kernel(const uint32_t* __restrict__ flags, uint32_t n,
                                                          uint32_t want, uint32_t* __restrict__ offs,
                                                          uint32_t* __restrict__ sums) {
    __shared__ uint32_t sh[1024];
    const uint32_t i = blockIdx.x * 1024u + threadIdx.x;
    const uint32_t f = (i < n && flags[i] == want) ? 1u : 0u;
    sh[threadIdx.x] = f;
    __syncthreads();
    for (uint32_t o = 1; o < 1024u; o <<= 1) {
        const uint32_t t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0u;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    if (i < n) offs[i] = sh[threadIdx.x] - f;
    if (threadIdx.x == 1023u) sums[blockIdx.x] = sh[1023];
}

__global__ __launch_bounds__(1024) void scan_sums_kernel(uint32_t* __restrict__ sums, uint32_t nchunks,
                                                         uint32_t* __restrict__ total) {
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    __shared__ uint32_t sh[1024];
    for (uint32_t base = 0; base < nchunks; base += 1024u) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < nchunks ? sums[i] : 0u;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t o = 1; o < 1024u; o <<= 1) {
            const uint32_t t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0u;
            __syncthreads();
            sh[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nchunks) sums[i] = carry + sh[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023u) carry += sh[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

// Per-child variant: flags has 8 entries per parent (child c of parent i at 8i + c); the
// selected children are written in parent order.
__global__ __launch_bounds__(256) void oct_children8_kernel(const uint32_t* __restrict__ blocks,
                                                            const uint32_t* __restrict__ flags, uint32_t want,
                                                            const uint32_t* __restrict__ offs,
                                                            const uint32_t* __restrict__ sums, uint32_t m,
                                                            uint32_t* __restrict__ children) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m || flags[i] != want) return;
    const uint32_t o = sums[i / 1024u] + offs[i];
    uint32_t x, y, z;
    unpack_block(blocks[i / 8u], x, y, z);
    const uint32_t c = i % 8u;
    const uint32_t cx = 2u * x + ((c >> 2) & 1u), cy = 2u * y + ((c >> 1) & 1u), cz = 2u * z + (c & 1u);
    children[o] = cx | (cy << 10) | (cz << 20);
}

hipError_t launch_oct_scan_children(const uint32_t* flags, uint32_t m, uint32_t want, uint32_t* offs, uint32_t* sums,
                                    uint32_t* total, const uint32_t* blocks, uint32_t* children,
                                    hipStream_t s) {
    if (m == 0) return hipMemsetAsync(total, 0, sizeof(uint32_t), s);
    const uint32_t chunks = (m + 1023u) / 1024u;
    hipLaunchKernelGGL(scan_chunk_kernel, dim3(chunks), dim3(1024), 0, s, flags, m, want, offs, sums);
    hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(1024), 0, s, sums, chunks, total);
    hipLaunchKernelGGL(oct_children8_kernel, dim3((m + 255u) / 256u), dim3(256), 0, s, blocks, flags, want, offs,
                       sums, m, children);
    return hipGetLastError();
}

// Every fine voxel of n blocks of edge B (packed coordinates in block units), written from
// out[0] in block order: the full-resolution fallback of low-confidence children.
__global__ __launch_bounds__(256) void oct_expand_kernel(const uint32_t* __restrict__ blocks, uint32_t n,
                                                         uint32_t B, uint32_t* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t cube = B * B * B;
    if (i >= (uint64_t)n * cube) return;
    const uint32_t b = (uint32_t)(i / cube), o = (uint32_t)(i % cube);
    uint32_t x, y, z;
    unpack_block(blocks[b], x, y, z);
    const uint32_t ox = o / (B * B), oy = (o / B) % B, oz = o % B;
    out[i] = (x * B + ox) | ((y * B + oy) << 10) | ((z * B + oz) << 20);
}

hipError_t launch_oct_expand(const uint32_t* blocks, uint32_t n, uint32_t B, uint32_t* out, hipStream_t s) {
    const uint64_t m = (uint64_t)n * B * B * B;
    if (m == 0) return hipSuccess;
    hipLaunchKernelGGL(oct_expand_kernel, dim3((uint32_t)((m + 255u) / 256u)), dim3(256), 0, s, blocks, n, B, out);
    return hipGetLastError();
}

// ------------------------------------------------------------------ finest level
// Exact single-tap values of a list of fine voxels (K per lane), scattered into the volume.
template <int K, bool FLAT>
__global__ __launch_bounds__(256) void oct_fine_kernel(OctFineArgs a) {
    const uint32_t base = (blockIdx.x * blockDim.x + threadIdx.x) * K;
    float vx[K], vy[K], vz[K], vz2[K];
    uint32_t ids[K];
#pragma unroll
    for (int r = 0; r < K; ++r) {
        const uint32_t idx = min(base + r, a.n - 1u);
        uint32_t x, y, z;
        unpack_block(a.voxels[idx], x, y, z);
        float v[3];
        const float fx = (float)x, fy = (float)y, fz = (float)z;
        for (int k = 0; k < 3; ++k) {
            float t = a.g.zero[k] + a.g.inc[k] * fx;
            t = t + a.g.inc[3 + k] * fy;
            v[k] = t + a.g.inc[6 + k] * fz;
        }
        vx[r] = v[0];
        vy[r] = v[1];
        vz[r] = v[2];
        const float cz = a.lz - v[2];
        vz2[r] = cz * cz;
        ids[r] = (x * a.g.n[1] + y) * a.g.n[2] + z;
    }
    const auto er = make_rsrc(a.E, a.e_bytes);
    using Off = ElemOffset<8>;
    const float top = Off::kBias + (float)a.Tc;
    float sum[K], comp[K], blk[K];
#pragma unroll
    for (int r = 0; r < K; ++r) sum[r] = comp[r] = blk[r] = 0.0f;
    uint32_t p = 0;
    while (p < a.P) {
        const uint32_t pend = min(p + a.kahan_block, a.P);
        // kUnroll pairs' loads in flight per lane; the block sums stay in pair order, so the
        // values equal the dense kernel's bit for bit
        constexpr int kUnroll = 4;
        // (flat wall: the z term is the voxel's; element offset from the biased quotient,
        // within 1/8 of q1 = fma(sqrt(s), inv2, w), the texture path's budget)
        auto tap = [&](uint32_t q, int r) {
            const float4 pr = a.pairs[q];
            float s;
            if (FLAT) {
                const float dx = pr.x - vx[r], dy = pr.y - vy[r];
                s = (dx * dx + dy * dy) + vz2[r];
            } else {
                s = sq3(pr.x - vx[r], pr.y - vy[r], pr.z - vz[r]);
            }
            const float y = __builtin_fmaf(__builtin_amdgcn_sqrtf(s), a.inv2, pr.w + Off::kBias);
            return FusedF32::pick(er, Off::clamped(y, top), q * a.lde * 8u, s);
        };
        for (; p + kUnroll <= pend; p += kUnroll) {
            float t[kUnroll][K];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u)
#pragma unroll
                for (int r = 0; r < K; ++r) t[u][r] = tap(p + u, r);
#pragma unroll
            for (int u = 0; u < kUnroll; ++u)
#pragma unroll
                for (int r = 0; r < K; ++r) blk[r] = blk[r] + t[u][r];
        }
        for (; p < pend; ++p)
#pragma unroll
            for (int r = 0; r < K; ++r) blk[r] = blk[r] + tap(p, r);
#pragma unroll
        for (int r = 0; r < K; ++r) {
            kahan_add(sum[r], comp[r], blk[r]);
            blk[r] = 0.0f;
        }
    }
#pragma unroll
    for (int r = 0; r < K; ++r)
        if (base + r < a.n) a.out[ids[r]] = kahan_result(sum[r], comp[r]);
}

// One voxel per lane up to 4M voxels (more lanes hide the gather latency of these sparse
// lists), two beyond (measured on 256^3 x 16384-pair captures: 1.1M voxels 10.8 ms with one
// per lane against 17.9 ms with two; 9.0M voxels 63 ms with two).
hipError_t launch_oct_fine(const OctFineArgs& a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    const int K = a.n < (1u << 22) ? 1 : 2;
    const uint32_t lanes = (a.n + K - 1) / K;
    const dim3 grid((lanes + 255u) / 256u);
    const bool flat = a.flat != 0u;
    if (K == 1) {
        if (flat) hipLaunchKernelGGL((oct_fine_kernel<1, true>), grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((oct_fine_kernel<1, false>), grid, dim3(256), 0, s, a);
    } else {
        if (flat) hipLaunchKernelGGL((oct_fine_kernel<2, true>), grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((oct_fine_kernel<2, false>), grid, dim3(256), 0, s, a);
    }
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void fill_kernel(float* __restrict__ p, uint64_t n, float v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

hipError_t launch_fill(float* p, uint64_t n, float v, hipStream_t s) {
    uint64_t blocks = (n + 255u) / 256u;
    if (blocks > 8192u) blocks = 8192u;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

// Dense list of every block of a level grid (row-major, z fastest).
__global__ __launch_bounds__(256) void oct_all_blocks_kernel(uint32_t nx, uint32_t ny, uint32_t nz,
                                                             uint32_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nx * ny * nz) return;
    const uint32_t z = i % nz, y = (i / nz) % ny, x = i / (nz * ny);
    out[i] = x | (y << 10) | (z << 20);
}

hipError_t launch_oct_all_blocks(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t* out, hipStream_t s) {
    const uint32_t n = nx * ny * nz;
    hipLaunchKernelGGL(oct_all_blocks_kernel, dim3((n + 255u) / 256u), dim3(256), 0, s, nx, ny, nz, out);
    return hipGetLastError();
}

}  // namespace bpx
