// gfx950 kernels of the NLOS backprojection: threshold tables, row staging, phasor FIR and
// the voxel gather-accumulate kernels.  Compiled with -ffp-contract=off (see bpx_device.hpp).
//
// Bin exactness.  The reference GPU path (backproject_cuda.cu:25-46, 144-162) assigns pair p
// and voxel v the single bin r = round(((lw + d1) + d2 + cw - t0) / deltaT) computed in fp32.
// For fixed pair, r is a monotone step function of the path length (and, for a confocal pair
// where d1 == d2, of the squared distance s).  The kernels therefore never divide: a table
// holds, for every bin j, the smallest input whose rounded bin is >= j (computed once per
// capture with the exact correctly-rounded formula, `min_true_key`), an approximate quotient
// picks the two candidate bins {m, m+1}, and one comparison against the table decides.  The
// host guarantees |approximate - exact quotient| < 0.5 (bpx_api.cpp, tap_error_bound), so the
// selected bin is bit-identical to the fp32 formula for every voxel and pair.
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "bpx_device.hpp"
#include "bpx_internal.hpp"

namespace bpx {

// =================================================================== threshold tables
// Confocal pair: input s = squared wall-point-to-voxel distance; d1 = d2 = sqrt(s).  The bin
// round((total - t0) / deltaT) is non-decreasing in total, so "bin >= j" is "total >= Z[j]"
// with Z[j] the smallest such total (ztable_kernel, kind 0): the search needs no division.
// total = ((lw + d) + d) + cw is non-decreasing in d and d = sqrt_rn(s) in s, so the smallest
// s with total >= Z[j] is the smallest s with sqrt_rn(s) >= d*, d* the smallest d with
// total(d) >= Z[j]: two short searches, the first over additions only, the second over one
// correctly rounded square root.
struct ChainPred {
    float lw, cw, z;
    __device__ bool operator()(uint32_t key) const {
        const float d = fkey_inv(key);
        float total = lw + d;
        total = total + d;
        total = total + cw;
        return total >= z;
    }
};
struct SqrtPred {
    float d;
    __device__ bool operator()(uint32_t key) const { return exact_sqrt(fkey_inv(key)) >= d; }
};

// The first search starts from a guess in double, d ~ (z - lw - cw) / 2, usually within a few
// floats of d* (three fp32 additions round total by at most ~1.5 ulp of total), so the
// straight-line bracket of 8 keys either side settles it; the general search keeps the result
// exact for the lanes where it does not (very small d against lw + cw).
//
// The second needs no search for d* in [2^-60, 2^60]: with d- the float below d* and
// m = (d- + d*) / 2 (exact in double: 26 significant bits), sqrt_rn(s) >= d* iff sqrt(s) > m, or
// sqrt(s) = m and the tie rounds to d* (round half to even: d*'s last mantissa bit is 0).  m^2
// has at most 52 significant bits, so it is exact in double, and sqrt(s) > m iff s > m^2.  The
// smallest such float is m^2 rounded up to a float, one float higher when m^2 is itself a float
// and the tie goes to d- (then sqrt_rn(m^2) = d-).  Outside that range (0, tiny, huge or
// infinite d*) the general search over s runs instead.
__device__ __forceinline__ float stable_threshold(float lw, float cw, float z) {
    const double gd = 0.5 * (((double)z - (double)lw) - (double)cw);
    const float g = gd > 0.0 ? (float)gd : 0.0f;
    const float dstar =
        fkey_inv(min_true_key_near(ChainPred{lw, cw, z}, fkey(g), fkey(0.0f), fkey(__builtin_inff())));
    // neighbours of non-negative finite floats by their bit patterns
    const float below = __uint_as_float(__float_as_uint(dstar) - 1u);
    const double m = 0.5 * ((double)dstar + (double)below);
    const double msq = m * m;
    float sg = (float)msq;
    sg = (double)sg < msq ? __uint_as_float(__float_as_uint(sg) + 1u) : sg;
    sg = (double)sg == msq && (__float_as_uint(dstar) & 1u) ? __uint_as_float(__float_as_uint(sg) + 1u) : sg;
    if (!(dstar >= 0x1p-60f && dstar <= 0x1p60f))   // 0, tiny, huge or infinite d*
        sg = fkey_inv(min_true_key(SqrtPred{dstar}, fkey(dstar * dstar), fkey(0.0f), fkey(__builtin_inff())));
    return sg;
}

// S[p][j] is written at float index p*pitch + j*stride + offset (stride 1, offset 0 for a
// plain table; the element stride/slot of a fused layout otherwise).
__global__ __launch_bounds__(256) void stable_kernel(const float* __restrict__ pair8, uint32_t P,
                                                     uint32_t Tc, const float* __restrict__ Z,
                                                     float* __restrict__ S, uint32_t pitch,
                                                     uint32_t stride, uint32_t offset) {
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = Tc + 1u;
    if (gid >= (uint64_t)P * n) return;
    const uint32_t p = (uint32_t)(gid / n), j = (uint32_t)(gid % n);
    S[(size_t)p * pitch + (size_t)j * stride + offset] = stable_threshold(pair8[8u * p + 7u], pair8[8u * p + 6u], Z[j]);
}

// Path-length space: kind 0 = single tap (backproject_cuda.cu:155), kind 1 / 2 = start / end
// of the CPU path's window, round(((t -/+ w/2) - t0)/deltaT) (backproject.hpp:88-89).
struct ZPred {
    int kind;
    float t0, dT, hw, jf;
    __device__ bool operator()(uint32_t key) const {
        const float t = fkey_inv(key);
        float y;
        if (kind == 0) y = t - t0;
        else if (kind == 1) y = (t - hw) - t0;
        else y = (t + hw) - t0;
        return round_away(exact_div(y, dT)) >= jf;
    }
};

__global__ __launch_bounds__(256) void ztable_kernel(int kind, float t0, float dT, float hw,
                                                     uint32_t n, float* __restrict__ Z) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const ZPred pred{kind, t0, dT, hw, (float)j};
    float g = t0 + ((float)j - 0.5f) * dT;
    if (kind == 1) g = g + hw;
    if (kind == 2) g = g - hw;
    const uint32_t k = min_true_key(pred, fkey(g), fkey(-__builtin_inff()), fkey(__builtin_inff()));
    Z[j] = fkey_inv(k);
}

hipError_t launch_stable(const float* pair8, uint32_t P, uint32_t Tc, const float* Z,
                         float* S, uint32_t pitch, uint32_t stride, uint32_t offset,
                         hipStream_t stream) {
    const uint64_t n = (uint64_t)P * (Tc + 1u);
    const uint32_t blocks = (uint32_t)((n + 255u) / 256u);
    hipLaunchKernelGGL(stable_kernel, dim3(blocks), dim3(256), 0, stream, pair8, P, Tc, Z, S, pitch,
                       stride, offset);
    return hipGetLastError();
}

hipError_t launch_ztable(int kind, float t0, float dT, float half_width, uint32_t n, float* Z,
                         hipStream_t stream) {
    hipLaunchKernelGGL(ztable_kernel, dim3((n + 255u) / 256u), dim3(256), 0, stream, kind, t0, dT,
                       half_width, n, Z);
    return hipGetLastError();
}

// =================================================================== row staging
__device__ __forceinline__ uint16_t f32_to_bf16_rne(float x) {
    const uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
    const uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(r >> 16);
}
__device__ __forceinline__ uint16_t f32_to_f16(float x) {
    return __builtin_bit_cast(uint16_t, (_Float16)x);
}

// Writes one sample (real: re only) into a padded destination row element.
__device__ __forceinline__ void store_sample(void* dst, size_t idx, int format, float re, float im) {
    switch (format) {
        case kF32: static_cast<float*>(dst)[idx] = re; break;
        case kF16: static_cast<uint16_t*>(dst)[idx] = f32_to_f16(re); break;
        case kBF16: static_cast<uint16_t*>(dst)[idx] = f32_to_bf16_rne(re); break;
        case kC64: static_cast<float2*>(dst)[idx] = make_float2(re, im); break;
        case kCF16:
            static_cast<uint32_t*>(dst)[idx] = (uint32_t)f32_to_f16(re) | ((uint32_t)f32_to_f16(im) << 16);
            break;
        default:
            static_cast<uint32_t*>(dst)[idx] =
                (uint32_t)f32_to_bf16_rne(re) | ((uint32_t)f32_to_bf16_rne(im) << 16);
            break;
    }
}

// dst[p][0] = dst[p][Tc+1..ldh) = 0, dst[p][1+b] = src[p][b] * inv_scale.  `src` holds real
// samples, or interleaved complex samples when the destination format is complex.
__global__ __launch_bounds__(256) void pad_rows_kernel(const float* __restrict__ src, uint32_t rows,
                                                       uint32_t Tc, uint32_t src_pitch, void* dst,
                                                       uint32_t ldh, int format, float inv_scale) {
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (uint64_t)rows * ldh) return;
    const uint32_t p = (uint32_t)(gid / ldh), i = (uint32_t)(gid % ldh);
    float re = 0.0f, im = 0.0f;
    if (i >= 1u && i <= Tc) {
        const uint32_t b = i - 1u;
        if (format_complex(format)) {
            re = src[2 * ((size_t)p * src_pitch + b)] * inv_scale;
            im = src[2 * ((size_t)p * src_pitch + b) + 1] * inv_scale;
        } else {
            re = src[(size_t)p * src_pitch + b] * inv_scale;
        }
    }
    store_sample(dst, gid, format, re, im);
}

// Sample slots of the fused layouts (see FusedF32 etc.): padded row index i holds raw sample
// i - 1; index 0 and indices past Tc hold 0.
__device__ __forceinline__ float padded_sample(const float* row, uint32_t i, uint32_t Tc, float inv_scale,
                                               int comp, int cplx) {
    if (i < 1u || i > Tc) return 0.0f;
    return cplx ? row[2 * (i - 1u) + comp] * inv_scale : row[i - 1u] * inv_scale;
}

// The sample slots of element j of a fused row of format F, before their conversion: v[0]
// (the real sample, or re), v[1] (im, or the next padded index's sample for the 16-bit real
// formats), v[2] and v[3] (re and im of the next padded index for the 16-bit complex formats).
template <int F>
__device__ __forceinline__ void fused_samples(const float* row, uint32_t j, uint32_t Tc, float inv_scale, float v[4]) {
    constexpr int cplx = (F == kC64 || F == kCF16 || F == kCBF16) ? 1 : 0;
    constexpr bool two = F != kF32 && F != kC64;   // 16-bit layouts hold two padded indices
    v[0] = padded_sample(row, j, Tc, inv_scale, 0, cplx);
    v[1] = cplx ? padded_sample(row, j, Tc, inv_scale, 1, 1) : two ? padded_sample(row, j + 1u, Tc, inv_scale, 0, 0) : 0.0f;
    v[2] = cplx && two ? padded_sample(row, j + 1u, Tc, inv_scale, 0, 1) : 0.0f;
    v[3] = cplx && two ? padded_sample(row, j + 1u, Tc, inv_scale, 1, 1) : 0.0f;
}

// Element e of a fused row of format F from its sample slots (fused_samples) and the
// threshold slot S.
template <int F>
__device__ __forceinline__ void write_fused(void* dst, uint64_t e, const float v[4], float S) {
    const uint32_t sb = __float_as_uint(S);
    if constexpr (F == kF32) {
        static_cast<float2*>(dst)[e] = make_float2(v[0], S);
    } else if constexpr (F == kF16 || F == kBF16) {
        const uint32_t ha = F == kF16 ? f32_to_f16(v[0]) : f32_to_bf16_rne(v[0]);
        const uint32_t hb = F == kF16 ? f32_to_f16(v[1]) : f32_to_bf16_rne(v[1]);
        static_cast<uint2*>(dst)[e] = make_uint2(sb, ha | (hb << 16));
    } else if constexpr (F == kC64) {
        static_cast<float4*>(dst)[e] = make_float4(v[0], v[1], S, 0.0f);
    } else {
        auto q = [](float x) -> uint32_t { return F == kCBF16 ? f32_to_bf16_rne(x) : f32_to_f16(x); };
        static_cast<uint4*>(dst)[e] = make_uint4(sb, q(v[0]) | (q(v[1]) << 16), q(v[2]) | (q(v[3]) << 16), 0u);
    }
}

// The job owning block `blk` of a pass: the last job whose first block is <= blk (jobs with no
// blocks share their first block with the next job and are skipped).  Uniform per block.
template <bool ZPASS>
__device__ __forceinline__ uint32_t stage_job_of(const StageJob* jobs, uint32_t n, uint32_t blk) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        const uint32_t f = ZPASS ? jobs[mid].z_first : jobs[mid].s_first;
        if (f <= blk) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Entry j of a job's path-length table Z (kind 0: single tap).
__device__ __forceinline__ void stage_z_entry(const StageJob& jb, uint32_t j) {
    const ZPred pred{0, jb.t0, jb.dT, 0.0f, (float)j};
    const float g = jb.t0 + ((float)j - 0.5f) * jb.dT;
    jb.Z[j] = fkey_inv(min_true_key(pred, fkey(g), fkey(-__builtin_inff()), fkey(__builtin_inff())));
}

// Row p of a job: the elements of its window (StageJob::win; the whole row without one) in
// chunks of kStageSpan, element base + threadIdx.x + 256 k of a chunk for k < kStageSpan / 256.
// The row, window and format come from block-uniform values (one body per format without
// format branches), and every lane issues the loads of all its elements of a chunk before the
// first threshold search.  Elements outside the window are never written: the plan clears
// its rows once when it is made, and no gather reads them.
template <int F>
__device__ __forceinline__ void stage_row_elements(const StageJob& jb, uint32_t p) {
    constexpr int kE = (int)(kStageSpan / 256u);
    constexpr bool cplx = F == kC64 || F == kCF16 || F == kCBF16;
    const float* row = jb.src + (size_t)p * jb.Tc * (cplx ? 2u : 1u);
    const float lw = jb.pair8[8u * p + 7u], cw = jb.pair8[8u * p + 6u];
    uint32_t lo = 0u, hi = jb.lde;
    if (jb.win) {
        lo = jb.win[2u * p];
        hi = jb.win[2u * p + 1u];
    }
    for (uint32_t base = lo; base < hi; base += kStageSpan) {
        float v[kE][4], z[kE];
#pragma unroll
        for (int k = 0; k < kE; ++k) {
            const uint32_t j = base + threadIdx.x + 256u * (uint32_t)k;
            fused_samples<F>(row, j, j < hi ? jb.Tc : 0u, jb.inv_scale, v[k]);
            z[k] = j < hi && j <= jb.Tc ? jb.Z[j] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < kE; ++k) {
            const uint32_t j = base + threadIdx.x + 256u * (uint32_t)k;
            if (j < hi) {
                const float S = j <= jb.Tc ? stable_threshold(lw, cw, z[k]) : 0.0f;
                write_fused<F>(jb.dst, (uint64_t)p * jb.lde + j, v[k], S);
            }
        }
    }
}

__device__ __forceinline__ void stage_row_block(const StageJob& jb, uint32_t p) {
    // rows of a plan whose samples are checked per execution: fp16 while the check allows it
    const int format = __builtin_amdgcn_readfirstlane(jb.fmt_flag && *jb.fmt_flag == 0u ? jb.alt_format : jb.format);
    switch (format) {
        case kF32: stage_row_elements<kF32>(jb, p); break;
        case kF16: stage_row_elements<kF16>(jb, p); break;
        case kBF16: stage_row_elements<kBF16>(jb, p); break;
        case kC64: stage_row_elements<kC64>(jb, p); break;
        case kCF16: stage_row_elements<kCF16>(jb, p); break;
        default: stage_row_elements<kCBF16>(jb, p); break;
    }
}

__global__ __launch_bounds__(256) void stage_z_kernel(const StageJob* __restrict__ jobs, uint32_t n_jobs,
                                                      uint32_t blocks) {
    for (uint32_t blk = blockIdx.x; blk < blocks; blk += gridDim.x) {
        const StageJob& jb = jobs[stage_job_of<true>(jobs, n_jobs, blk)];
        const uint32_t j = (blk - jb.z_first) * 256u + threadIdx.x;
        if (j <= jb.Tc) stage_z_entry(jb, j);
    }
}

// Each work-group stages runs of kStageRun consecutive rows (row blocks of the pass, numbered
// over all jobs): one job search per run (the chain of dependent job-table loads is paid once),
// then the next job's first row is checked as the run crosses into it (jobs with no rows are
// stepped over, as in the search).
constexpr uint32_t kStageRun = 2;
__global__ __launch_bounds__(256) void stage_rows_kernel(const StageJob* __restrict__ jobs, uint32_t n_jobs,
                                                         uint32_t blocks) {
    for (uint32_t run = blockIdx.x * kStageRun; run < blocks; run += gridDim.x * kStageRun) {
        uint32_t ji = stage_job_of<false>(jobs, n_jobs, run);
        const uint32_t end = min(run + kStageRun, blocks);
        for (uint32_t blk = run; blk < end; ++blk) {
            while (ji + 1u < n_jobs && jobs[ji + 1u].s_first <= blk) ++ji;
            const StageJob& jb = jobs[ji];
            stage_row_block(jb, blk - jb.s_first);
        }
    }
}

__global__ __launch_bounds__(256) void crop_rows_kernel(const float* __restrict__ src, uint32_t src_T,
                                                        const uint32_t* __restrict__ map, uint32_t rows,
                                                        uint32_t Tc, uint32_t from, uint32_t cfrom,
                                                        uint32_t ncopy, float* __restrict__ raw) {
    const uint64_t n = (uint64_t)rows * Tc, stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; gid < n; gid += stride) {
        const uint32_t r = (uint32_t)(gid / Tc), c = (uint32_t)(gid % Tc);
        float v = 0.0f;
        if (c >= cfrom && c - cfrom < ncopy) v = src[(size_t)(map ? map[r] : r) * src_T + from + (c - cfrom)];
        raw[gid] = v;
    }
}

hipError_t launch_crop_rows(const float* src, uint32_t src_T, const uint32_t* map, uint32_t rows, uint32_t Tc,
                            uint32_t from, uint32_t cfrom, uint32_t ncopy, float* raw, hipStream_t stream) {
    const uint64_t n = (uint64_t)rows * Tc;
    if (n == 0) return hipSuccess;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255u) / 256u, 1u << 20);
    hipLaunchKernelGGL(crop_rows_kernel, dim3(blocks), dim3(256), 0, stream, src, src_T, map, rows, Tc, from, cfrom,
                       ncopy, raw);
    return hipGetLastError();
}

uint32_t stage_z_blocks(const StageJob& j) { return (j.Tc + 1u + 255u) / 256u; }
uint32_t stage_row_blocks(const StageJob& j) {
    return j.rows;   // one row block per row
}

hipError_t launch_stage(const StageJob* jobs, uint32_t n_jobs, uint32_t z_blocks, uint32_t row_blocks,
                        hipStream_t stream) {
    if (n_jobs == 0) return hipSuccess;
    constexpr uint32_t kMaxGrid = 1u << 20;
    if (z_blocks)
        hipLaunchKernelGGL(stage_z_kernel, dim3(std::min(z_blocks, kMaxGrid)), dim3(256), 0, stream, jobs, n_jobs,
                           z_blocks);
    if (row_blocks)
        hipLaunchKernelGGL(stage_rows_kernel, dim3(std::min((row_blocks + kStageRun - 1u) / kStageRun, kMaxGrid)),
                           dim3(256), 0, stream, jobs, n_jobs, row_blocks);
    return hipGetLastError();
}

hipError_t launch_pad_rows(const float* src, uint32_t rows, uint32_t Tc, uint32_t src_pitch,
                           void* dst, uint32_t ldh, int format, float inv_scale,
                           hipStream_t stream) {
    const uint64_t n = (uint64_t)rows * ldh;
    hipLaunchKernelGGL(pad_rows_kernel, dim3((uint32_t)((n + 255u) / 256u)), dim3(256), 0, stream,
                       src, rows, Tc, src_pitch, dst, ldh, format, inv_scale);
    return hipGetLastError();
}

// Exclusive prefix sums in double, one wave per row: PS[p][k] = sum_{i<k} H[p][i], k < T.
__global__ __launch_bounds__(256) void prefix_kernel(const float* __restrict__ src, uint32_t rows,
                                                     uint32_t T, int cplx, void* dst, uint32_t ldp) {
    const uint32_t row = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    if (row >= rows) return;
    double carry_re = 0.0, carry_im = 0.0;
    for (uint32_t c0 = 0; c0 < T; c0 += 64u) {
        const uint32_t i = c0 + lane;
        double xr = 0.0, xi = 0.0;
        if (i < T) {
            if (cplx) {
                xr = src[2 * ((size_t)row * T + i)];
                xi = src[2 * ((size_t)row * T + i) + 1];
            } else {
                xr = src[(size_t)row * T + i];
            }
        }
        // inclusive wave scan
        double sr = xr, si = xi;
        for (int off = 1; off < 64; off <<= 1) {
            const double ur = __shfl_up(sr, off, 64);
            const double ui = __shfl_up(si, off, 64);
            if ((int)lane >= off) {
                sr += ur;
                si += ui;
            }
        }
        if (i < T) {
            const double er = carry_re + sr - xr, ei = carry_im + si - xi;
            if (cplx) static_cast<double2*>(dst)[(size_t)row * ldp + i] = make_double2(er, ei);
            else static_cast<double*>(dst)[(size_t)row * ldp + i] = er;
        }
        carry_re += __shfl(sr, 63, 64);
        carry_im += __shfl(si, 63, 64);
    }
}

hipError_t launch_prefix(const float* src, uint32_t rows, uint32_t T, int complex_samples,
                         void* dst, uint32_t ldp, hipStream_t stream) {
    hipLaunchKernelGGL(prefix_kernel, dim3((rows + 3u) / 4u), dim3(256), 0, stream, src, rows, T,
                       complex_samples, dst, ldp);
    return hipGetLastError();
}

// Phasor-field filter (backproject.hpp:459-476): out[t] = sum_{j<pulse} in[t+j-pulse/2] *
// (c_j + i s_j) over 0 < idx < Tc, products rounded, summed in j order; written straight
// into a complex destination row of ldh elements, out[t] at element lead + t (lead 1: the
// padded rows of the gather kernels; lead 0: plain [rows][Tc] complex rows), zero elsewhere.
__global__ __launch_bounds__(256) void phasor_fir_kernel(const float* __restrict__ src, uint32_t rows,
                                                         uint32_t Tc, const float* __restrict__ tc,
                                                         const float* __restrict__ ts, int pulse,
                                                         void* dst, uint32_t ldh, uint32_t lead, int format,
                                                         float inv_scale) {
    extern __shared__ float taps[];
    for (int j = threadIdx.x; j < pulse; j += blockDim.x) {
        taps[2 * j] = tc[j];
        taps[2 * j + 1] = ts[j];
    }
    __syncthreads();
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (uint64_t)rows * ldh) return;
    const uint32_t p = (uint32_t)(gid / ldh), i = (uint32_t)(gid % ldh);
    float re = 0.0f, im = 0.0f;
    if (i >= lead && i - lead < Tc) {
        const int64_t t = (int64_t)(i - lead);
        const float* row = src + (size_t)p * Tc;
        const int half = pulse / 2;
        for (int j = 0; j < pulse; ++j) {
            const int64_t idx = t + j - half;
            if (idx > 0 && idx < (int64_t)Tc) {
                const float x = row[idx];
                re = re + x * taps[2 * j];
                im = im + x * taps[2 * j + 1];
            }
        }
        re = re * inv_scale;
        im = im * inv_scale;
    }
    store_sample(dst, gid, format, re, im);
}

hipError_t launch_phasor_fir(const float* src, uint32_t rows, uint32_t Tc, const float* taps_c,
                             const float* taps_s, int pulse, void* dst, uint32_t ldh, uint32_t lead,
                             int format, float inv_scale, hipStream_t stream) {
    const uint64_t n = (uint64_t)rows * ldh;
    if (n == 0) return hipSuccess;
    if ((n + 255u) / 256u > 0x7fffffffull || (uint64_t)lead + Tc > ldh) return hipErrorInvalidValue;
    hipLaunchKernelGGL(phasor_fir_kernel, dim3((uint32_t)((n + 255u) / 256u)), dim3(256),
                       sizeof(float) * 2 * pulse, stream, src, rows, Tc, taps_c, taps_s, pulse, dst,
                       ldh, lead, format, inv_scale);
    return hipGetLastError();
}

// Plain streaming copy (16 B per lane, grid-stride): the HBM bandwidth reference and the
// calibration kernel of the FETCH_SIZE / WRITE_SIZE counters.
__global__ __launch_bounds__(256) void copy_kernel(float4* __restrict__ dst, const float4* __restrict__ src,
                                                   uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

hipError_t launch_copy(void* dst, const void* src, uint64_t bytes, hipStream_t stream) {
    const uint64_t n = bytes / 16u;
    hipLaunchKernelGGL(copy_kernel, dim3(256u * 32u), dim3(256), 0, stream, static_cast<float4*>(dst),
                       static_cast<const float4*>(src), n);
    return hipGetLastError();
}

// max |x| over n floats, as the bit pattern of a non-negative float (atomic max on uint).
__global__ __launch_bounds__(256) void abs_max_kernel(const float* __restrict__ src, uint64_t n,
                                                      unsigned int* result) {
    float m = 0.0f;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const float a = fabsf(src[i]);
        m = (a > m || a != a) ? a : m;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const float o = __shfl_xor(m, off, 64);
        m = (o > m || o != o) ? o : m;
    }
    if ((threadIdx.x & 63u) == 0u) atomicMax(result, __float_as_uint(m));
}

hipError_t launch_abs_max(const float* src, uint64_t n, float* result_dev, hipStream_t stream) {
    uint64_t blocks = (n + 255u) / 256u;
    if (blocks > 4096u) blocks = 4096u;
    if (blocks == 0u) blocks = 1u;
    hipLaunchKernelGGL(abs_max_kernel, dim3((uint32_t)blocks), dim3(256), 0, stream, src, n,
                       reinterpret_cast<unsigned int*>(result_dev));
    return hipGetLastError();
}

// Magnitude and phase of a complex volume (interleaved re, im): |v| = sqrt(re^2 + im^2) with
// the squares and the sum in double and one rounding to float (the double-precision hypot the
// host's std::abs(std::complex<float>) evaluates; infinite if either part is), phase =
// atan2(im, re) in double rounded to float.  Either output may be null.
__global__ __launch_bounds__(256) void complex_post_kernel(const float2* __restrict__ v, uint64_t n,
                                                           float* __restrict__ mag, float* __restrict__ phase) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const float2 c = v[i];
        const double re = (double)c.x, im = (double)c.y;
        if (mag) {
            const bool inf = __builtin_isinf(c.x) || __builtin_isinf(c.y);
            mag[i] = inf ? __builtin_inff() : (float)__builtin_sqrt(re * re + im * im);
        }
        if (phase) phase[i] = (float)atan2(im, re);
    }
}

hipError_t launch_complex_post(const void* complex_volume, uint64_t n, float* mag, float* phase,
                               hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint64_t blocks = std::min<uint64_t>((n + 255u) / 256u, 1u << 16);
    hipLaunchKernelGGL(complex_post_kernel, dim3((uint32_t)blocks), dim3(256), 0, stream,
                       static_cast<const float2*>(complex_volume), n, mag, phase);
    return hipGetLastError();
}

// Whether every sample is an integer that fp16 holds exactly (|x| <= 2048): one flag word,
// set by any wave that finds a sample that is not (NaN and infinities included).
__device__ __forceinline__ bool exact16_sample(float x) {
    return __builtin_fabsf(x) <= 2048.0f && x == __builtin_rintf(x);
}

// Sets *flag when a sample is not an integer of magnitude <= 2048: 16-byte loads over the
// aligned body (src 16-byte aligned), single samples for the tail.
__global__ __launch_bounds__(256) void exact16_check_kernel(const float* __restrict__ src, uint64_t n,
                                                            uint32_t* flag) {
    bool bad = false;
    const uint64_t n4 = n / 4u, stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const float4* s4 = reinterpret_cast<const float4*>(src);
    for (uint64_t i = tid; i < n4; i += stride) {
        const float4 v = s4[i];
        bad = bad || !(exact16_sample(v.x) && exact16_sample(v.y) && exact16_sample(v.z) && exact16_sample(v.w));
    }
    for (uint64_t i = 4u * n4 + tid; i < n; i += stride) bad = bad || !exact16_sample(src[i]);
    if (__ballot(bad) != 0ull && (threadIdx.x & 63u) == 0u) atomicOr(flag, 1u);
}

hipError_t launch_exact16_check(const float* src, uint64_t n, uint32_t* flag, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    if (reinterpret_cast<uintptr_t>(src) % 16u) return hipErrorInvalidValue;
    const uint64_t blocks = std::min<uint64_t>((n / 4u + 255u) / 256u + 1u, 8192u);
    hipLaunchKernelGGL(exact16_check_kernel, dim3((uint32_t)blocks), dim3(256), 0, stream, src, n, flag);
    return hipGetLastError();
}

// =================================================================== gather kernels
Tiles make_tiles(const uint32_t n[3], int K) {
    Tiles t;
    t.tx = (n[0] + 7u) / 8u;
    t.ty = (n[1] + 7u) / 8u;
    t.tz = (n[2] + 4u * K - 1u) / (4u * K);
    t.count = t.tx * t.ty * t.tz;
    return t;
}

// Lane -> voxel mapping of a 256-lane work-group covering 8 x 8 x 4K voxels.
struct LaneVoxels {
    uint32_t ix, iy, iz0;
};
// Wave w (0..3) of a tile covers the 4 x 4 x 4K quarter (w & 1, w >> 1) of its 8 x 8 columns.
__device__ __forceinline__ LaneVoxels lane_voxels_of(const Tiles& t, int K, uint32_t tile, uint32_t w) {
    const uint32_t tz = tile % t.tz, txy = tile / t.tz;
    const uint32_t ty = txy % t.ty, tx = txy / t.ty;
    const uint32_t lane = threadIdx.x & 63u;
    LaneVoxels lv;
    lv.ix = tx * 8u + (w & 1u) * 4u + (lane & 3u);
    lv.iy = ty * 8u + (w >> 1) * 4u + ((lane >> 2) & 3u);
    lv.iz0 = tz * (4u * (uint32_t)K) + (lane >> 4);
    return lv;
}
__device__ __forceinline__ LaneVoxels lane_voxels(const Tiles& t, int K) {
    return lane_voxels_of(t, K, xcd_remap(blockIdx.x, gridDim.x), threadIdx.x >> 6);
}

// A.1 voxel position, backproject_cuda.cu:144-148.
__device__ __forceinline__ void voxel_position(const Grid& g, uint32_t x, uint32_t y, uint32_t z,
                                               float v[3]) {
    const float fx = (float)x, fy = (float)y, fz = (float)z;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float a = g.zero[k] + g.inc[k] * fx;
        a = a + g.inc[3 + k] * fy;
        v[k] = a + g.inc[6 + k] * fz;
    }
}

template <class T>
__device__ __forceinline__ void store_voxel(void* out, size_t id, T val);
template <>
__device__ __forceinline__ void store_voxel<float>(void* out, size_t id, float val) {
    static_cast<float*>(out)[id] = val;
}
template <>
__device__ __forceinline__ void store_voxel<float2>(void* out, size_t id, float2 val) {
    static_cast<float2*>(out)[id] = val;
}
__device__ __forceinline__ float scaled(float v, float s) { return v * s; }
__device__ __forceinline__ float2 scaled(float2 v, float s) { return make_float2(v.x * s, v.y * s); }

// ------------------------------------------------------------------- confocal single tap
// Every pair has laser point == camera point, so d1 == d2 == sqrt(s) with
// s = ((dx*dx + dy*dy) + dz*dz).  Per voxel-pair update: s (one add when the wall is flat),
// one hardware sqrt for the approximate quotient, a clamp, one table read, one compare and
// the sample read.  SEP: voxels of a lane share x and y (inc[6] == inc[7] == 0).
// FLAT: every wall point has the same z.
// Sample sources of the confocal kernel.  Split: threshold table and padded rows are two
// arrays (two dependent loads).  Fused: one interleaved array (one load, see FusedF32).
template <class Smp>
struct SplitSource {
    using acc_t = typename Smp::acc_t;
    __amdgpu_buffer_rsrc_t hr, sr;
    uint32_t ldh, lds, soff_h = 0, soff_s = 0;
    __device__ explicit SplitSource(const ConfocalArgs& a)
        : hr(make_rsrc(a.H, a.h_bytes)), sr(make_rsrc(a.S, a.s_bytes)), ldh(a.ldh), lds(a.lds) {}
    __device__ __forceinline__ void set_pair(uint32_t p) {
        soff_s = p * lds * 4u;
        soff_h = p * ldh * (uint32_t)Smp::kBytes;
    }
    __device__ __forceinline__ acc_t fetch(uint32_t j, float s) const {
        const float th = buf_f32(sr, j * 4u, soff_s);
        const uint32_t hj = j + (s >= th ? 1u : 0u);
        return Smp::load(hr, hj, soff_h);
    }
    // quotients are not biased: y = q, top = Tc
    static constexpr float kBias = 0.0f;
    __device__ __forceinline__ void accumulate(acc_t& acc, float y, float top, float s) const {
        acc_add(acc, fetch((uint32_t)(int)__builtin_amdgcn_fmed3f(y, 0.0f, top), s));
    }
};
template <class Fz>
struct FusedSource {
    using acc_t = typename Fz::acc_t;
    __amdgpu_buffer_rsrc_t er;
    uint32_t lde, soff = 0;
    __device__ explicit FusedSource(const ConfocalArgs& a) : er(make_rsrc(a.H, a.h_bytes)), lde(a.ldh) {}
    __device__ __forceinline__ void set_pair(uint32_t p) { soff = p * lde * (uint32_t)Fz::kElemBytes; }
    // quotients biased for ElemOffset: y = q + kBias, top = kBias + Tc
    using Off = ElemOffset<(uint32_t)Fz::kElemBytes>;
    static constexpr float kBias = Off::kBias;
    __device__ __forceinline__ void accumulate(acc_t& acc, float y, float top, float s) const {
        pick_accumulate<Fz>(acc, er, Off::clamped(y, top), soff, s, 0);
    }
};

// Wave quarter w of one 8 x 8 x 4K tile of a confocal capture (the body of the gather
// kernels; no work-group barrier, so a wave can run it alone).
template <class Src, int K, bool SEP, bool FLAT>
__device__ __forceinline__ void confocal_tile(const ConfocalArgs& a, uint32_t tile, uint32_t w) {
    using acc_t = typename Src::acc_t;
    const LaneVoxels lv = lane_voxels_of(a.t, K, tile, w);
    const uint32_t cx = min(lv.ix, a.g.n[0] - 1u), cy = min(lv.iy, a.g.n[1] - 1u);
    float vx[SEP ? 1 : K], vy[SEP ? 1 : K], vz[K], dz2[K];
#pragma unroll
    for (int r = 0; r < K; ++r) {
        const uint32_t cz = min(lv.iz0 + 4u * r, a.g.n[2] - 1u);
        float v[3];
        voxel_position(a.g, cx, cy, cz, v);
        if (SEP) {
            if (r == 0) {
                vx[0] = v[0];
                vy[0] = v[1];
            }
        } else {
            vx[r] = v[0];
            vy[r] = v[1];
        }
        vz[r] = v[2];
        const float dz = a.lz - v[2];
        dz2[r] = dz * dz;
    }
    Src src(a);
    // sources of fused rows take quotients biased by Src::kBias (exact: integers < 2^24)
    const float top = Src::kBias + (float)a.Tc;
    acc_t sum[K], comp[K], blk[K];
#pragma unroll
    for (int r = 0; r < K; ++r) {
        acc_zero(sum[r]);
        acc_zero(comp[r]);
        acc_zero(blk[r]);
    }
    uint32_t p = 0;
    while (p < a.P) {
        const uint32_t pend = min(p + a.kahan_block, a.P);
        for (; p < pend; ++p) {
            const float4 pr = a.pairs[p];
            src.set_pair(p);
            const float w = Src::kBias == 0.0f ? pr.w : pr.w + Src::kBias;
            float axy = 0.0f;
            if (SEP) {
                const float dx = pr.x - vx[0], dy = pr.y - vy[0];
                axy = dx * dx + dy * dy;
            }
#pragma unroll
            for (int r = 0; r < K; ++r) {
                float s;
                if (SEP) {
                    if (FLAT) {
                        s = axy + dz2[r];
                    } else {
                        const float dz = pr.z - vz[r];
                        s = axy + dz * dz;
                    }
                } else {
                    s = sq3(pr.x - vx[r], pr.y - vy[r], pr.z - vz[r]);
                }
                const float d = __builtin_amdgcn_sqrtf(s);
                src.accumulate(blk[r], __builtin_fmaf(d, a.inv2, w), top, s);
            }
        }
#pragma unroll
        for (int r = 0; r < K; ++r) {
            kahan_add(sum[r], comp[r], blk[r]);
            acc_zero(blk[r]);
        }
    }
    if (lv.ix < a.g.n[0] && lv.iy < a.g.n[1]) {
#pragma unroll
        for (int r = 0; r < K; ++r) {
            const uint32_t iz = lv.iz0 + 4u * r;
            if (iz < a.g.n[2]) {
                const size_t id = ((size_t)lv.ix * a.g.n[1] + lv.iy) * a.g.n[2] + iz;
                store_voxel(a.out, id, scaled(kahan_result(sum[r], comp[r]), a.scale));
            }
        }
    }
}

template <class Src, int K, bool SEP, bool FLAT>
__global__ __launch_bounds__(256) void confocal_kernel(ConfocalArgs a) {
    confocal_tile<Src, K, SEP, FLAT>(a, xcd_remap(blockIdx.x, gridDim.x), threadIdx.x >> 6);
}

// ------------------------------------------------------------------- LDS-staged confocal gather
// Fused rows of any storage format, separable grid, flat wall.  A 256-lane work-group owns a 16 x 16 x KZ voxel
// tile (lane = one (x, y) column, KZ voxels along z).  For every pair, the bin guesses of all
// the tile's voxels lie in a short window [jlo, jhi]: the tile's bounding box gives the nearest
// and farthest squared distances, and the window is widened by 2 bins on each side, far more
// than the approximate quotient's error.  The window's fused elements are copied to LDS with
// coalesced loads and the voxel lookups read LDS instead of the texture path.  Four pairs per
// stage (one per wave), double-buffered: the loads for stage b + 1 are in flight while the
// waves gather stage b.
//
// The guesses are not clamped to [0, Tc]: elements outside the row (index < 0 or > Tc + 1) are
// staged as zeros.  A guess j picks sample j or j - 1; a guess below 0 (quotient <= -1) or
// above Tc (quotient >= Tc + 1) can only stand for a bin outside the capture, whose value is 0,
// and inside [0, Tc] the guess equals the clamped one.  So every update adds the same value as
// in confocal_kernel, in the same pair order with the same Kahan blocks: the volumes are
// bit-identical.
constexpr uint32_t kSegTileXY = 16;   // tile edge in x and y
// staged elements per pair: 512 of 8 B, or 384 of 16 B (complex64)
__host__ __device__ constexpr uint32_t seg_elems(int elem_bytes) { return elem_bytes == 8 ? 512u : 384u; }

// One tile of the LDS-staged gather for the whole work-group (NW waves).  It ends with a
// work-group barrier, so a persistent caller may run the next tile right after it.  BOUNDED:
// stage only the 64-element groups a window touches (pays off for short windows, costs
// branches for long ones).
// Sample formats that stage one component of complex rows (FusedC64Part).
template <class T, class = void>
struct is_component : std::false_type {};
template <class T>
struct is_component<T, std::void_t<decltype(T::kComponent)>> : std::true_type {};

template <class Fz, int KZ, int NW, bool BOUNDED = false, bool UNROLL = true>
__device__ __forceinline__ void lds_tile(const ConfocalArgs& a, uint32_t tile) {
    using acc_t = typename Fz::acc_t;
    // EG: bytes of a row element in HBM; E: bytes staged per element (its leading part)
    constexpr uint32_t EG = Fz::kElemBytes, E = Fz::kLdsBytes, SEG = seg_elems(Fz::kLdsBytes), NL = SEG / 64u;
    // NW waves (= pairs per stage); tile depth TZ: NW / 4 z-slabs of KZ voxels per lane
    constexpr uint32_t TZ = KZ * NW / 4;
    using elem_t = std::conditional_t<E == 8, uint2, uint4>;
    __shared__ elem_t seg[2][NW][SEG];
    __shared__ int meta[2][NW][2];   // window start, staged element count
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63u;
    const uint32_t n0 = a.g.n[0], n1 = a.g.n[1], n2 = a.g.n[2];
    const uint32_t ntz = (n2 + TZ - 1u) / TZ, nty = (n1 + kSegTileXY - 1u) / kSegTileXY;
    const uint32_t tz = tile % ntz, txy = tile / ntz, ty = txy % nty, tx = txy / nty;
    const uint32_t ix = tx * kSegTileXY + (w & 1u) * 8u + (lane & 7u);
    const uint32_t iy = ty * kSegTileXY + ((w >> 1) & 1u) * 8u + (lane >> 3);
    const uint32_t iz0 = tz * TZ + (w >> 2) * KZ;
    const uint32_t cx = min(ix, n0 - 1u), cy = min(iy, n1 - 1u);
    float vx = 0.0f, vy = 0.0f, dz2[KZ];
#pragma unroll
    for (int r = 0; r < KZ; ++r) {
        float v[3];
        voxel_position(a.g, cx, cy, min(iz0 + r, n2 - 1u), v);
        if (r == 0) {
            vx = v[0];
            vy = v[1];
        }
        const float dz = a.lz - v[2];
        dz2[r] = dz * dz;
    }
    // bounding box of the tile's voxel positions: each coordinate is a sum of terms monotone
    // in one index each (rounding is monotone), so its extremes are at the 8 corners
    const uint32_t x0 = min(tx * kSegTileXY, n0 - 1u), x1 = min(tx * kSegTileXY + kSegTileXY - 1u, n0 - 1u);
    const uint32_t y0 = min(ty * kSegTileXY, n1 - 1u), y1 = min(ty * kSegTileXY + kSegTileXY - 1u, n1 - 1u);
    const uint32_t z0 = min(tz * TZ, n2 - 1u), z1 = min(tz * TZ + TZ - 1u, n2 - 1u);
    float bmin[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
    float bmax[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        float v[3];
        voxel_position(a.g, (c & 1) ? x1 : x0, (c & 2) ? y1 : y0, (c & 4) ? z1 : z0, v);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            bmin[k] = fminf(bmin[k], v[k]);
            bmax[k] = fmaxf(bmax[k], v[k]);
        }
    }
    const int last_elem = (int)a.Tc + 1;   // sample Tc + 1 (padded index) = 0 ends every row
    const uint32_t row_bytes = (a.Tc + 2u) * EG;
    const __amdgpu_buffer_rsrc_t er = make_rsrc(a.H, a.h_bytes);
    // the wall is flat: the z terms of the box distances are the same for every pair
    const float ez = a.lz - fminf(fmaxf(a.lz, bmin[2]), bmax[2]);
    const float fz = fmaxf(fabsf(a.lz - bmin[2]), fabsf(a.lz - bmax[2]));
    const float ez2 = ez * ez, fz2 = fz * fz;

    auto load = [&](__amdgpu_buffer_rsrc_t r, int off, int soff) {
        if constexpr (is_component<Fz>::value) {
            return Fz::stage_load(r, off, soff);
        } else if constexpr (E == 8) {
            const auto t = __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, 0);
            return make_uint2(t[0], t[1]);
        } else {
            const auto t = __builtin_amdgcn_raw_buffer_load_b128(r, off, soff, 0);
            return make_uint4(t[0], t[1], t[2], t[3]);
        }
    };
    // Stage registers of this wave's pair: window and NL elements per lane.
    elem_t pre[NL];
    int pjlo = 0, pcnt = 0;
    auto fetch = [&](uint32_t p0) {
        const uint32_t p = p0 + w;
        pjlo = 0;
        pcnt = 0;
        if (p < a.P) {
            const float4 pr = a.pairs[p];
            const float ex = pr.x - fminf(fmaxf(pr.x, bmin[0]), bmax[0]);
            const float ey = pr.y - fminf(fmaxf(pr.y, bmin[1]), bmax[1]);
            const float fx = fmaxf(fabsf(pr.x - bmin[0]), fabsf(pr.x - bmax[0]));
            const float fy = fmaxf(fabsf(pr.y - bmin[1]), fabsf(pr.y - bmax[1]));
            const float smin = (ex * ex + ey * ey) + ez2;
            const float smax = (fx * fx + fy * fy) + fz2;
            const int jl = (int)__builtin_fmaf(__builtin_amdgcn_sqrtf(smin), a.inv2, pr.w) - 2;
            const int jh = (int)__builtin_fmaf(__builtin_amdgcn_sqrtf(smax), a.inv2, pr.w) + 2;
            pjlo = __builtin_amdgcn_readfirstlane(jl);
            pcnt = __builtin_amdgcn_readfirstlane(min(jh + 2 - jl, (int)SEG));   // elements jl .. jh + 1
        }
        if (p < a.P && pjlo >= 0) {
            // Window inside or past the row: a descriptor over row p alone returns 0 for every
            // element past its end, so no lane needs a bounds test (unbounded: elements past the
            // window are loaded too and never read).
            const __amdgpu_buffer_rsrc_t rr =
                make_rsrc(static_cast<const char*>(a.H) + (size_t)p * a.ldh * EG, row_bytes);
            const int vo = (pjlo + (int)lane) * (int)EG;
#pragma unroll
            for (uint32_t m = 0; m < NL; ++m)
                if (!BOUNDED || (int)(64u * m) < pcnt) pre[m] = load(rr, vo + (int)(64u * EG * m), 0);
        } else {
            // window starting before the row (or no pair): elements below 0 are zeros
            const uint32_t soff = p < a.P ? p * a.ldh * EG : 0u;
#pragma unroll
            for (uint32_t m = 0; m < NL; ++m) {
                const int e = (int)(lane + 64u * m), idx = pjlo + e;
                elem_t v{};
                if (e < pcnt && idx >= 0 && idx <= last_elem) v = load(er, idx * (int)EG, (int)soff);
                pre[m] = v;
            }
        }
    };
    auto commit = [&](uint32_t buf) {
        // (bounded: only the groups the window touches; no lookup reads past element pcnt - 1)
#pragma unroll
        for (uint32_t m = 0; m < NL; ++m)
            if (!BOUNDED || (int)(64u * m) < pcnt) seg[buf][w][lane + 64u * m] = pre[m];
        if (lane == 0) {
            meta[buf][w][0] = pjlo;
            meta[buf][w][1] = pcnt;
        }
    };

    acc_t sum[KZ], comp[KZ], blk[KZ];
#pragma unroll
    for (int r = 0; r < KZ; ++r) {
        acc_zero(sum[r]);
        acc_zero(comp[r]);
        acc_zero(blk[r]);
    }
    const uint32_t nstage = (a.P + NW - 1u) / NW;
    const char* segb = reinterpret_cast<const char*>(&seg[0][0][0]);
    uint32_t in_block = 0;
    fetch(0u);
    commit(0u);
    __syncthreads();
    // One pair of a stage: slot = its LDS segment, start = its window start (wave-uniform).
    // The LDS byte offset of element j of a pair's window is E (j - start + slot SEG).  The
    // update takes it from the mantissa of y = fma(sqrt(s), inv2, W), W = w + (slot SEG - start)
    // + ElemOffset<E>::kBias: the offset of element floor(y - kBias) = floor(q) - start +
    // slot SEG for the approximate quotient q = fma(sqrt(s), inv2, w).  W and y each round to
    // 1/E, so floor(q) comes from a quotient within 1/E of the one the host bounds (bound
    // < 1/4, 1/E <= 1/8: total < 1/2, which keeps the element index in {bin, bin + 1}, the two
    // the threshold pick decides between).  The window (elements start .. ) begins two
    // elements below the lowest such index, so y - kBias >= slot SEG + 1 > 0.
    using Off = ElemOffset<E>;
    auto pair_step = [&](const float4 pr, uint32_t slot, uint32_t start) {
        const float W = pr.w + (float)(int)(slot * SEG - start + (uint32_t)Off::kBias);
        const float dx = pr.x - vx, dy = pr.y - vy;
        const float axy = dx * dx + dy * dy;
#pragma unroll
        for (int r = 0; r < KZ; ++r) {
            const float s = axy + dz2[r];
            const float y = __builtin_fmaf(__builtin_amdgcn_sqrtf(s), a.inv2, W);
            lds_accumulate<Fz>(blk[r], segb + Off::of(y), s, 0);
        }
        if (++in_block == a.kahan_block) {
#pragma unroll
            for (int r = 0; r < KZ; ++r) {
                kahan_add(sum[r], comp[r], blk[r]);
                acc_zero(blk[r]);
            }
            in_block = 0;
        }
    };
    for (uint32_t b = 0; b < nstage; ++b) {
        const uint32_t buf = b & 1u;
        if (b + 1u < nstage) fetch((b + 1u) * NW);
        const uint32_t p0 = b * NW;
        // Full stages with 8 real voxels per lane: the stage's window starts and pair data read
        // once into scalar registers and the pair loop unrolled.  Deeper tiles (16 voxels per
        // lane) and complex accumulators keep the loop: unrolled, they need more registers than
        // 4 waves per SIMD allow.
        constexpr bool kUnrollStage = UNROLL && KZ == 8 && std::is_same<acc_t, float>::value;
        if (kUnrollStage && p0 + NW <= a.P) {
            uint32_t start[NW];
            float4 prs[NW];
#pragma unroll
            for (uint32_t i = 0; i < NW; ++i) {
                start[i] = __builtin_amdgcn_readfirstlane((uint32_t)meta[buf][i][0]);
                prs[i] = a.pairs[p0 + i];
            }
#pragma unroll
            for (uint32_t i = 0; i < NW; ++i) pair_step(prs[i], buf * NW + i, start[i]);
        } else {
            const uint32_t np = min((uint32_t)NW, a.P - p0);
            for (uint32_t i = 0; i < np; ++i)
                pair_step(a.pairs[p0 + i], buf * NW + i, __builtin_amdgcn_readfirstlane((uint32_t)meta[buf][i][0]));
        }
        if (b + 1u < nstage) commit(buf ^ 1u);
        __syncthreads();
    }
    if (in_block) {
#pragma unroll
        for (int r = 0; r < KZ; ++r) kahan_add(sum[r], comp[r], blk[r]);
    }
    if (ix < n0 && iy < n1) {
#pragma unroll
        for (int r = 0; r < KZ; ++r) {
            const uint32_t iz = iz0 + r;
            if (iz < n2) {
                const size_t id = ((size_t)ix * n1 + iy) * n2 + iz;
                if constexpr (is_component<Fz>::value)   // one component of a complex volume
                    static_cast<float*>(a.out)[2 * id + Fz::kComponent] = kahan_result(sum[r], comp[r]) * a.scale;
                else
                    store_voxel(a.out, id, scaled(kahan_result(sum[r], comp[r]), a.scale));
            }
        }
    }
}

template <class Fz, int KZ, int NW>
__global__ __launch_bounds__(NW * 64) void confocal_lds_kernel(ConfocalArgs a) {
    // rows staged in the other format this execution: the other instantiation runs instead
    if (a.fmt_flag && __builtin_amdgcn_readfirstlane(*a.fmt_flag) != a.fmt_want) return;
    lds_tile<Fz, KZ, NW>(a, xcd_remap(blockIdx.x, gridDim.x));
}

// ------------------------------------------------------------------- batched captures
// One persistent launch serves every item of a batch of confocal captures from a single queue:
//   stage   -- kBatchStageBlocks fused rows of one capture
//   LDS     -- one 16 x 16 x KZ tile of a capture whose bin windows fit the LDS segment
//   texture -- one 8 x 8 x 4K tile of any other capture (the 4 waves take its quarters)
// Work-groups take items in queue order (one atomic by thread 0, broadcast through LDS between
// two barriers, so every wave of a work-group sees the same item and they leave together).  A
// capture's gather items wait until its staging items are done; the queue holds every staging
// item first (in capture order) and then the gathers in the same order, so the first captures'
// gathers start while the last captures are still being staged, and a wait is rare.  Staging
// items are large (64K row elements: one atomic per item on the shared queue head).  Staging
// items are taken before the gathers that wait for them and never wait themselves, so every
// wait ends; it is also bounded (spin_limit polls, then the
// error word is set and the work-group goes on), and the item loop is bounded by the item
// count, so the launch always drains.  Each tile runs the single-capture kernels' tile body:
// every volume is bit-identical to its own plan's.  The work-group finishing a capture's last
// gather item stamps the device wall clock (per-capture latency).
// A struct every lane reads from the same address, moved word by word into scalar registers:
// the buffer descriptors built from it are then wave-uniform (no per-pair waterfall loops).
template <class T>
__device__ __forceinline__ T wave_uniform_load(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "word-sized struct");
    T out;
    const uint32_t* src = reinterpret_cast<const uint32_t*>(p);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
    return out;
}

// KINDS: the item kinds the instantiation serves (bit 0 staging, bit 1 LDS-staged, bit 2 texture
// path); an LDS-only queue compiles without the other paths and needs fewer registers.  NW:
// waves per work-group -- 4, or 8 for an LDS-only queue of 16^3 tiles with 8 voxels per lane
// (the single-capture 512-lane kernel's shape and unrolled stage loop).
template <class Fz, class Src, int K, int KZ, int KINDS, int NW = 4>
__global__ __launch_bounds__(NW * 64) void batch_queue_kernel(BatchQueueArgs b) {
    static_assert(NW == 4 || KINDS == 2, "8-wave work-groups serve LDS-staged tiles only");
    __shared__ uint32_t s_item;
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (uint32_t taken = 0; taken <= b.n_items; ++taken) {
        if (threadIdx.x == 0) s_item = atomicAdd(b.next, 1u);
        __syncthreads();
        const uint32_t it = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (it >= b.n_items) return;
        const uint2 item = b.items[it];
        const uint32_t c = __builtin_amdgcn_readfirstlane(item.x & 0x0fffffffu);
        const uint32_t type = __builtin_amdgcn_readfirstlane(item.x >> 28);
        const uint32_t y = __builtin_amdgcn_readfirstlane(item.y);
        if ((KINDS & 1) && type == kBatchItemStage) {
            const StageJob jb = wave_uniform_load(b.jobs + c);
            const uint32_t total = jb.rows;
            const uint32_t end = min(y + kBatchStageBlocks, total);
            for (uint32_t blk = y; blk < end; ++blk) stage_row_block(jb, blk);
            __threadfence();
            __syncthreads();
            if (threadIdx.x == 0) atomicSub(&b.staged[c], 1u);
            continue;
        }
        if ((KINDS & 1) && threadIdx.x == 0) {
            uint32_t polls = 0;
            while (__hip_atomic_load(&b.staged[c], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
                if (++polls > b.spin_limit) {
                    atomicOr(b.error, 1u);
                    break;
                }
                __builtin_amdgcn_s_sleep(8);
            }
        }
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");   // the staged rows, for every wave
        const ConfocalArgs a = wave_uniform_load(b.caps + c);
        if ((KINDS & 2) && type == kBatchItemLds) {
            // 4 waves: rolled stage loop (fewer registers, more work-groups)
            lds_tile<Fz, KZ, NW, true, NW == 8>(a, y);
        } else if (KINDS & 4) {
            const uint32_t f = __builtin_amdgcn_readfirstlane(b.flags[c]);
            if (f == 3u) confocal_tile<Src, K, true, true>(a, y, w);
            else if (f == 1u) confocal_tile<Src, K, true, false>(a, y, w);
            else confocal_tile<Src, K, false, false>(a, y, w);
        }
        __threadfence();
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t left = atomicSub(&b.remaining[c], 1u);
            if (left == 1u)
                atomicMax(reinterpret_cast<unsigned long long*>(&b.done[c]), (unsigned long long)wall_clock64());
        }
    }
}

// Texture-path tiles served by WAVES, for a batch run as two concurrent persistent launches (the
// LDS-staged tiles in batch_queue_kernel<.., KINDS = 2>, these beside it on a second stream, the
// rows staged beforehand): every wave takes the next (capture, tile * 4 + quarter) item.  The queue
// head and the completion counters are updated by every lane with an increment that is 1 on lane
// 0 and 0 elsewhere, never under a lane-divergent branch (a branch on "lane 0" around the head
// update lets the compiler split the wave into lane 0 and the rest, which then re-read a stale
// head and never leave); there is no work-group barrier, and the item loop is bounded by the
// item count.
template <int K, class Src>
__global__ __launch_bounds__(256) void batch_tex_kernel(BatchQueueArgs b) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t one = lane == 0u ? 1u : 0u;
    for (uint32_t taken = 0; taken <= b.n_items; ++taken) {
        // all lanes active: the first active lane is lane 0, whose old value is the item
        const uint32_t it = __builtin_amdgcn_readfirstlane(atomicAdd(b.next, one));
        if (it >= b.n_items) return;
        const uint2 item = b.items[it];
        const uint32_t c = __builtin_amdgcn_readfirstlane(item.x & 0x0fffffffu);
        const uint32_t q = __builtin_amdgcn_readfirstlane(item.y);  // tile * 4 + quarter
        const ConfocalArgs a = wave_uniform_load(b.caps + c);
        const uint32_t f = __builtin_amdgcn_readfirstlane(b.flags[c]);
        if (f == 3u) confocal_tile<Src, K, true, true>(a, q >> 2, q & 3u);
        else if (f == 1u) confocal_tile<Src, K, true, false>(a, q >> 2, q & 3u);
        else confocal_tile<Src, K, false, false>(a, q >> 2, q & 3u);
        __threadfence();
        const uint32_t left = atomicSub(&b.remaining[c], one);
        const unsigned long long stamp = (one && left == 1u) ? (unsigned long long)wall_clock64() : 0ull;
        atomicMax(reinterpret_cast<unsigned long long*>(&b.done[c]), stamp);
    }
}

// Before the queue: every capture's path-length table (blocks < z_blocks, as stage_z_kernel)
// and, in the last block, the queue reset: both heads and the error word 0, per-capture
// counters from their initial values, completion stamps 0, the start stamp.
__global__ __launch_bounds__(256) void batch_prologue_kernel(const StageJob* __restrict__ jobs, uint32_t n_jobs,
                                                             uint32_t z_blocks, BatchQueueArgs b,
                                                             const uint32_t* __restrict__ init, uint32_t n,
                                                             uint64_t* start) {
    if (blockIdx.x < z_blocks) {
        const StageJob& jb = jobs[stage_job_of<true>(jobs, n_jobs, blockIdx.x)];
        const uint32_t j = (blockIdx.x - jb.z_first) * 256u + threadIdx.x;
        if (j <= jb.Tc) stage_z_entry(jb, j);
        return;
    }
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        b.staged[i] = init[2u * i];
        b.remaining[i] = init[2u * i + 1u];
        b.done[i] = 0ull;
    }
    if (threadIdx.x == 0) {
        b.next[0] = 0u;   // the queue head; the texture-path and 16^3-tile heads of the
        b.next[1] = 0u;   // two- and three-launch forms
        b.next[2] = 0u;
        *b.error = 0u;
        *start = wall_clock64();
    }
}

template <class Fz, class Src, int K, int KZ, int KINDS, int NW = 4>
static const void* batch_queue_fn() {
    return reinterpret_cast<const void*>(batch_queue_kernel<Fz, Src, K, KZ, KINDS, NW>);
}
// the 16^3-tile queue of 512-lane work-groups (real formats)
static const void* batch_wide_of(int format) {
    switch (format) {
        case kF32: return batch_queue_fn<FusedF32, FusedSource<FusedF32>, 2, 8, 2, 8>();
        case kF16: return batch_queue_fn<FusedH16<false>, FusedSource<FusedH16<false>>, 2, 8, 2, 8>();
        case kBF16: return batch_queue_fn<FusedH16<true>, FusedSource<FusedH16<true>>, 2, 8, 2, 8>();
        default: return nullptr;
    }
}
template <int K, int KZ, int KINDS>
static const void* batch_queue_of(int format) {
    switch (format) {
        case kF32: return batch_queue_fn<FusedF32, FusedSource<FusedF32>, K, KZ, KINDS>();
        case kF16: return batch_queue_fn<FusedH16<false>, FusedSource<FusedH16<false>>, K, KZ, KINDS>();
        case kBF16: return batch_queue_fn<FusedH16<true>, FusedSource<FusedH16<true>>, K, KZ, KINDS>();
        case kC64: return batch_queue_fn<FusedC64, FusedSource<FusedC64>, K, KZ, KINDS>();
        case kCF16: return batch_queue_fn<FusedC16<false>, FusedSource<FusedC16<false>>, K, KZ, KINDS>();
        case kCBF16: return batch_queue_fn<FusedC16<true>, FusedSource<FusedC16<true>>, K, KZ, KINDS>();
        default: return nullptr;
    }
}
// launches 1: every kind in one kernel; 2 and 3: the (16 x 16 x kz) LDS-staged tiles alone
// (K unused; 3 keeps depth 8 beside the 16^3-tile queue)
static const void* batch_queue(int K, int kz, int format, int launches) {
    if (format_complex(format) || launches == 3) kz = 8;   // complex: two accumulators per voxel
    if (launches >= 2) return kz == 16 ? batch_queue_of<2, 16, 2>(format) : batch_queue_of<2, 8, 2>(format);
    if (K == 2) return kz == 16 ? batch_queue_of<2, 16, 7>(format) : batch_queue_of<2, 8, 7>(format);
    if (K == 4) return kz == 16 ? batch_queue_of<4, 16, 7>(format) : batch_queue_of<4, 8, 7>(format);
    return nullptr;
}
template <int K>
static const void* batch_tex_of(int format) {
    switch (format) {
        case kF32: return reinterpret_cast<const void*>(batch_tex_kernel<K, FusedSource<FusedF32>>);
        case kF16: return reinterpret_cast<const void*>(batch_tex_kernel<K, FusedSource<FusedH16<false>>>);
        case kBF16: return reinterpret_cast<const void*>(batch_tex_kernel<K, FusedSource<FusedH16<true>>>);
        case kC64: return reinterpret_cast<const void*>(batch_tex_kernel<K, FusedSource<FusedC64>>);
        case kCF16: return reinterpret_cast<const void*>(batch_tex_kernel<K, FusedSource<FusedC16<false>>>);
        case kCBF16: return reinterpret_cast<const void*>(batch_tex_kernel<K, FusedSource<FusedC16<true>>>);
        default: return nullptr;
    }
}
static const void* batch_tex(int K, int format) {
    return K == 4 ? batch_tex_of<4>(format) : K == 2 ? batch_tex_of<2>(format) : nullptr;
}

static int resident_of(const void* fn, int threads = 256) {
    int dev = 0, cus = 0, per_cu = 0;
    if (!fn || hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, 0) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return cus * per_cu;
}

int batch_queue_resident_groups(int K, int kz, int format, int launches) {
    return resident_of(batch_queue(K, kz, format, launches));
}
int batch_tex_resident_groups(int K, int format) { return resident_of(batch_tex(K, format)); }
int batch_wide_resident_groups(int format) { return resident_of(batch_wide_of(format), 512); }
hipError_t launch_batch_wide(const BatchQueueArgs& b, int format, uint32_t groups, hipStream_t s) {
    const void* fn = batch_wide_of(format);
    if (!fn || groups == 0u) return hipErrorInvalidValue;
    void* args[] = {const_cast<BatchQueueArgs*>(&b)};
    return hipLaunchKernel(fn, dim3(groups), dim3(512), args, 0, s);
}

hipError_t launch_batch_tex(const BatchQueueArgs& b, int K, int format, uint32_t groups, hipStream_t s) {
    const void* fn = batch_tex(K, format);
    if (!fn || groups == 0u) return hipErrorInvalidValue;
    void* args[] = {const_cast<BatchQueueArgs*>(&b)};
    return hipLaunchKernel(fn, dim3(groups), dim3(256), args, 0, s);
}
int batch_lds_depth(int kz, int format) { return format_complex(format) ? 8 : kz; }

hipError_t launch_batch_queue(const BatchQueueArgs& b, int K, int kz, int format, int launches, uint32_t groups,
                              hipStream_t s) {
    const void* fn = batch_queue(K, kz, format, launches);
    if (!fn || groups == 0u) return hipErrorInvalidValue;
    void* args[] = {const_cast<BatchQueueArgs*>(&b)};
    return hipLaunchKernel(fn, dim3(groups), dim3(256), args, 0, s);
}

hipError_t launch_batch_prologue(const StageJob* jobs, uint32_t n_jobs, uint32_t z_blocks, const BatchQueueArgs& b,
                                 const uint32_t* init, uint32_t n, uint64_t* start, hipStream_t s) {
    hipLaunchKernelGGL(batch_prologue_kernel, dim3(z_blocks + 1u), dim3(256), 0, s, jobs, n_jobs, z_blocks, b, init,
                       n, start);
    return hipGetLastError();
}

bool batch_lds_applies(const Grid& g, float inv2, int format, bool sep, bool flat, int kz) {
    // the batch kernel stages complex64 rows as 16-B elements (both components in one pass)
    const uint32_t seg = seg_elems(format == kC64 ? FusedC64::kLdsBytes : 8);
    return sep && flat && confocal_lds_fits(g.inc, inv2, kz, seg);
}
uint32_t batch_lds_tiles(const Grid& g, int kz) {
    return ((g.n[0] + kSegTileXY - 1u) / kSegTileXY) * ((g.n[1] + kSegTileXY - 1u) / kSegTileXY) *
           ((g.n[2] + (uint32_t)kz - 1u) / (uint32_t)kz);
}

// ------------------------------------------------------------------- exhaustive grid
// Pairs p = l*C + c of an L-laser x C-SPAD capture.  d1 (laser leg) is computed once per
// (voxel, laser) and d2 (SPAD leg) once per (voxel, SPAD) per CS-chunk, both with the exact
// correctly-rounded sqrt; the total path ((lw + d1) + d2) + cw is then exact and its bin is
// decided against the path-length thresholds Z held in LDS.
template <class Smp, int K, int CS>
__global__ __launch_bounds__(256) void grid_kernel(GridArgs a) {
    using acc_t = typename Smp::acc_t;
    extern __shared__ float zs[];
    for (uint32_t j = threadIdx.x; j <= a.Tc; j += blockDim.x) zs[j] = a.Z[j];
    __syncthreads();
    const LaneVoxels lv = lane_voxels(a.t, K);
    const uint32_t cx = min(lv.ix, a.g.n[0] - 1u), cy = min(lv.iy, a.g.n[1] - 1u);
    float v[K][3];
#pragma unroll
    for (int r = 0; r < K; ++r) voxel_position(a.g, cx, cy, min(lv.iz0 + 4u * r, a.g.n[2] - 1u), v[r]);
    const auto hr = make_rsrc(a.H, a.h_bytes);
    const float fTc = (float)a.Tc;
    acc_t sum[K], comp[K], blk[K];
#pragma unroll
    for (int r = 0; r < K; ++r) {
        acc_zero(sum[r]);
        acc_zero(comp[r]);
        acc_zero(blk[r]);
    }
    uint32_t since = 0;
    for (uint32_t c0 = 0; c0 < a.C; c0 += CS) {
        const uint32_t ncs = min((uint32_t)CS, a.C - c0);
        float d2[CS][K], cw[CS];
#pragma unroll
        for (int cs = 0; cs < CS; ++cs) {
            const float4 sp = a.spads[min(c0 + cs, a.C - 1u)];
            cw[cs] = sp.w;
#pragma unroll
            for (int r = 0; r < K; ++r)
                d2[cs][r] = exact_sqrt(sq3(v[r][0] - sp.x, v[r][1] - sp.y, v[r][2] - sp.z));
        }
        for (uint32_t l = 0; l < a.L; ++l) {
            const float4 lp = a.lasers[l];
            float d1[K];
#pragma unroll
            for (int r = 0; r < K; ++r)
                d1[r] = exact_sqrt(sq3(lp.x - v[r][0], lp.y - v[r][1], lp.z - v[r][2]));
            // all CS x K gathers of this laser are issued before any is added (a partial last
            // chunk reads its last SPAD's row for the missing ones and drops those values)
            acc_t val[CS][K];
#pragma unroll
            for (int cs = 0; cs < CS; ++cs) {
                const uint32_t p = l * a.C + min(c0 + cs, a.C - 1u);
                const uint32_t soff_h = p * a.ldh * (uint32_t)Smp::kBytes;
#pragma unroll
                for (int r = 0; r < K; ++r) {
                    float total = lp.w + d1[r];
                    total = total + d2[cs][r];
                    total = total + cw[cs];
                    const float q1 = __builtin_fmaf(total, a.inv, a.c1);
                    const uint32_t j = (uint32_t)(int)__builtin_amdgcn_fmed3f(q1, 0.0f, fTc);
                    const uint32_t hj = j + (total >= zs[j] ? 1u : 0u);
                    val[cs][r] = Smp::load(hr, hj, soff_h);
                }
            }
            if (ncs == (uint32_t)CS) {
#pragma unroll
                for (int cs = 0; cs < CS; ++cs)
#pragma unroll
                    for (int r = 0; r < K; ++r) acc_add(blk[r], val[cs][r]);
            } else {
#pragma unroll
                for (int cs = 0; cs < CS; ++cs)
                    if ((uint32_t)cs < ncs)
#pragma unroll
                        for (int r = 0; r < K; ++r) acc_add(blk[r], val[cs][r]);
            }
            since += ncs;
            if (since >= a.kahan_block) {
                since = 0;
#pragma unroll
                for (int r = 0; r < K; ++r) {
                    kahan_add(sum[r], comp[r], blk[r]);
                    acc_zero(blk[r]);
                }
            }
        }
    }
    if (lv.ix < a.g.n[0] && lv.iy < a.g.n[1]) {
#pragma unroll
        for (int r = 0; r < K; ++r) {
            kahan_add(sum[r], comp[r], blk[r]);
            const uint32_t iz = lv.iz0 + 4u * r;
            if (iz < a.g.n[2]) {
                const size_t id = ((size_t)lv.ix * a.g.n[1] + lv.iy) * a.g.n[2] + iz;
                store_voxel(a.out, id, scaled(kahan_result(sum[r], comp[r]), a.scale));
            }
        }
    }
}

// ------------------------------------------------------------------- arbitrary pairs
// RAW: the samples are the caller's unpadded rows [P][Tc] (ldh = Tc) instead of padded rows;
// padded index hj is raw sample hj - 1, and hj = 0 or Tc + 1 (outside the row) reads 0.
template <class Smp, int K, bool RAW = false>
__global__ __launch_bounds__(256) void pairs_kernel(PairsArgs a) {
    using acc_t = typename Smp::acc_t;
    extern __shared__ float zs[];
    for (uint32_t j = threadIdx.x; j <= a.Tc; j += blockDim.x) zs[j] = a.Z[j];
    __syncthreads();
    const LaneVoxels lv = lane_voxels(a.t, K);
    const uint32_t cx = min(lv.ix, a.g.n[0] - 1u), cy = min(lv.iy, a.g.n[1] - 1u);
    float v[K][3];
#pragma unroll
    for (int r = 0; r < K; ++r) voxel_position(a.g, cx, cy, min(lv.iz0 + 4u * r, a.g.n[2] - 1u), v[r]);
    const auto hr = make_rsrc(a.H, a.h_bytes);
    const float fTc = (float)a.Tc;
    acc_t sum[K], comp[K], blk[K];
#pragma unroll
    for (int r = 0; r < K; ++r) {
        acc_zero(sum[r]);
        acc_zero(comp[r]);
        acc_zero(blk[r]);
    }
    // the K gathers of a pair are issued before any is added
    auto fetch = [&](uint32_t q, acc_t* val) {
        const float* pr = a.pairs + 8u * q;
        const float cpx = pr[0], cpy = pr[1], cpz = pr[2];
        const float lpx = pr[3], lpy = pr[4], lpz = pr[5];
        const float cw = pr[6], lw = pr[7];
        const uint32_t soff_h = q * a.ldh * (uint32_t)Smp::kBytes;
#pragma unroll
        for (int r = 0; r < K; ++r) {
            const float d1 = exact_sqrt(sq3(lpx - v[r][0], lpy - v[r][1], lpz - v[r][2]));
            const float d2 = exact_sqrt(sq3(v[r][0] - cpx, v[r][1] - cpy, v[r][2] - cpz));
            float total = lw + d1;
            total = total + d2;
            total = total + cw;
            const float q1 = __builtin_fmaf(total, a.inv, a.c1);
            const uint32_t j = (uint32_t)(int)__builtin_amdgcn_fmed3f(q1, 0.0f, fTc);
            const uint32_t hj = j + (total >= zs[j] ? 1u : 0u);
            if constexpr (RAW) {
                const uint32_t b = hj - 1u;   // wraps to 2^32 - 1 for hj = 0
                const bool in_row = b < a.Tc;
                const acc_t x = Smp::load(hr, in_row ? b : 0u, soff_h);
                val[r] = in_row ? x : acc_t{};
            } else {
                val[r] = Smp::load(hr, hj, soff_h);
            }
        }
    };
    uint32_t p = 0;
    while (p < a.P) {
        const uint32_t pend = min(p + a.kahan_block, a.P);
        for (; p + 2u <= pend; p += 2u) {   // two pairs' gathers in flight; sums in pair order
            acc_t v0[K], v1[K];
            fetch(p, v0);
            fetch(p + 1u, v1);
#pragma unroll
            for (int r = 0; r < K; ++r) acc_add(blk[r], v0[r]);
#pragma unroll
            for (int r = 0; r < K; ++r) acc_add(blk[r], v1[r]);
        }
        for (; p < pend; ++p) {
            acc_t v0[K];
            fetch(p, v0);
#pragma unroll
            for (int r = 0; r < K; ++r) acc_add(blk[r], v0[r]);
        }
#pragma unroll
        for (int r = 0; r < K; ++r) {
            kahan_add(sum[r], comp[r], blk[r]);
            acc_zero(blk[r]);
        }
    }
    if (lv.ix < a.g.n[0] && lv.iy < a.g.n[1]) {
#pragma unroll
        for (int r = 0; r < K; ++r) {
            const uint32_t iz = lv.iz0 + 4u * r;
            if (iz < a.g.n[2]) {
                const size_t id = ((size_t)lv.ix * a.g.n[1] + lv.iy) * a.g.n[2] + iz;
                store_voxel(a.out, id, scaled(kahan_result(sum[r], comp[r]), a.scale));
            }
        }
    }
}

// ------------------------------------------------------------------- window (CPU path)
// backproject.hpp:75-96 on the CPU grid (octree_volume.hpp:147-167): float differences,
// double squares/sum/sqrt (:42-53), wvw = float(d1 + d2), total = (lw + wvw) + cw; window
// [max(0, round(((t - w/2) - t0)/dT)), min(T-1, round(((t + w/2) - t0)/dT))) summed through
// double prefix sums.
__device__ __forceinline__ double a2_dist(float ax, float ay, float az, float bx, float by, float bz) {
    double x = (double)(ax - bx), y = (double)(ay - by), z = (double)(az - bz);
    x = x * x;
    y = y * y;
    z = z * z;
    return __builtin_sqrt((x + y) + z);
}

// clamp(round-bin(t), 0, T-1) from an approximate quotient and the thresholds Z.
__device__ __forceinline__ uint32_t clamped_bin(float t, float inv, float c1, const float* Z, float hi) {
    const float q1 = __builtin_fmaf(t, inv, c1);
    const uint32_t j = (uint32_t)(int)__builtin_amdgcn_fmed3f(q1, 1.0f, hi);
    return j - 1u + (t >= Z[j] ? 1u : 0u);
}

// CPU-path semantics, one voxel per lane.  The window-edge tables sit in LDS when they fit
// (ZLDS); two pairs are processed per iteration so that their edge lookups and prefix-sum
// loads are in flight together (the sums still run in pair order).
constexpr uint32_t kWindowLdsBins = 6144;   // 2 x 24 KB of window-edge thresholds

template <bool CPLX, bool ZLDS>
__global__ __launch_bounds__(256) void window_kernel(WindowArgs a) {
    __shared__ float zl[ZLDS ? kWindowLdsBins : 1], zh[ZLDS ? kWindowLdsBins : 1];
    if (ZLDS) {
        for (uint32_t j = threadIdx.x; j < a.T; j += blockDim.x) {
            zl[j] = a.Zlo[j];
            zh[j] = a.Zhi[j];
        }
        __syncthreads();
    }
    const float* Zlo = ZLDS ? zl : a.Zlo;
    const float* Zhi = ZLDS ? zh : a.Zhi;
    const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t nvox = (uint64_t)a.n[0] * a.n[1] * a.n[2];
    if (id >= nvox) return;
    const uint32_t z = (uint32_t)(id % a.n[2]);
    const uint32_t y = (uint32_t)((id / a.n[2]) % a.n[1]);
    const uint32_t x = (uint32_t)(id / ((uint64_t)a.n[2] * a.n[1]));
    const float vx = a.zero[0] + (float)x * a.vsize[0];
    const float vy = a.zero[1] + (float)y * a.vsize[1];
    const float vz = a.zero[2] + (float)z * a.vsize[2];
    const float hiT = (float)(a.T - 1u);
    double sre = 0.0, sim = 0.0;
    using elem_t = typename std::conditional<CPLX, double2, double>::type;
    const elem_t* ps = static_cast<const elem_t*>(a.PS);
    auto window = [&](uint32_t p, uint32_t& first, uint32_t& last) {
        const float* pr = a.pairs + 8u * p;
        const double d1 = a2_dist(pr[3], pr[4], pr[5], vx, vy, vz);
        const double d2 = a.confocal ? d1 : a2_dist(vx, vy, vz, pr[0], pr[1], pr[2]);
        const float wvw = (float)(d1 + d2);
        float total = pr[7] + wvw;
        total = total + pr[6];
        first = clamped_bin(total, a.inv, a.c1_lo, Zlo, hiT);
        last = clamped_bin(total, a.inv, a.c1_hi, Zhi, hiT);
    };
    auto add = [&](uint32_t p, uint32_t first, uint32_t last, const elem_t& f, const elem_t& l) {
        (void)p;
        if (last > first) {
            if constexpr (CPLX) {
                sre += l.x - f.x;
                sim += l.y - f.y;
            } else {
                sre += l - f;
            }
        }
    };
    constexpr uint32_t U = 2;   // measured: 2 pairs 103 ms, 4 pairs 165 ms at 128^3 (1 pair: 204 ms)
    uint32_t p = 0;
    for (; p + U <= a.P; p += U) {
        uint32_t f[U], l[U];
        elem_t lo[U], hi[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) window(p + u, f[u], l[u]);
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            const elem_t* r = ps + (size_t)(p + u) * a.ldp;
            lo[u] = r[f[u]];
            hi[u] = r[l[u]];
        }
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) add(p + u, f[u], l[u], lo[u], hi[u]);
    }
    for (; p < a.P; ++p) {
        uint32_t f0, l0;
        window(p, f0, l0);
        const elem_t* r0 = ps + (size_t)p * a.ldp;
        add(p, f0, l0, r0[f0], r0[l0]);
    }
    if (CPLX) static_cast<float2*>(a.out)[id] = make_float2((float)sre, (float)sim);
    else static_cast<float*>(a.out)[id] = (float)sre;
}

hipError_t launch_window(const WindowArgs& a, int complex_samples, hipStream_t stream) {
    const uint64_t nvox = (uint64_t)a.n[0] * a.n[1] * a.n[2];
    const uint32_t blocks = (uint32_t)((nvox + 255u) / 256u);
    const bool lds = a.T <= kWindowLdsBins;
    if (complex_samples) {
        if (lds) hipLaunchKernelGGL((window_kernel<true, true>), dim3(blocks), dim3(256), 0, stream, a);
        else hipLaunchKernelGGL((window_kernel<true, false>), dim3(blocks), dim3(256), 0, stream, a);
    } else {
        if (lds) hipLaunchKernelGGL((window_kernel<false, true>), dim3(blocks), dim3(256), 0, stream, a);
        else hipLaunchKernelGGL((window_kernel<false, false>), dim3(blocks), dim3(256), 0, stream, a);
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------- dispatch
// Variants v = 3*layout + k: layout 0 = split tables, 1 = fused layout; k: 0 -> K = 4,
// 1 -> K = 8, 2 -> K = 2 voxels per lane.  Variants 6, 7 and 8: the LDS-staged kernel with
// 16^3 tiles of 256 lanes, 16 x 16 x 8 tiles of 256 lanes, 16^3 tiles of 512 lanes (complex
// 16-bit rows: 16 x 16 x 8 for 6 and 7) for fused rows on a separable grid and a flat wall when
// the bin windows fit, otherwise the fused layout with K = 4 / 8 / 2.
int variant_k(int variant) {
    switch (variant % 3) {
        case 1: return 8;
        case 2: return 2;
        default: return 4;
    }
}
bool variant_fused(int variant) { return variant >= 3; }
int variant_count() { return 9; }
int fused_elem_bytes(int format) {
    switch (format) {
        case kF32: case kF16: case kBF16: return 8;
        default: return 16;
    }
}
int fused_s_offset(int format) {
    switch (format) {
        case kF32: return FusedF32::kSOffset;
        case kC64: return FusedC64::kSOffset;
        default: return 0;
    }
}

template <class Src, int K>
static hipError_t confocal_dispatch_k(const ConfocalArgs& a, bool sep, bool flat, hipStream_t s) {
    const dim3 grid(a.t.count), block(256);
    if (sep && flat) hipLaunchKernelGGL((confocal_kernel<Src, K, true, true>), grid, block, 0, s, a);
    else if (sep) hipLaunchKernelGGL((confocal_kernel<Src, K, true, false>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((confocal_kernel<Src, K, false, false>), grid, block, 0, s, a);
    return hipGetLastError();
}

template <class Src>
static hipError_t confocal_dispatch(const ConfocalArgs& a, int variant, bool sep, bool flat,
                                    hipStream_t s) {
    switch (variant_k(variant)) {
        case 8: return confocal_dispatch_k<Src, 8>(a, sep, flat, s);
        case 2: return confocal_dispatch_k<Src, 2>(a, sep, flat, s);
        default: return confocal_dispatch_k<Src, 4>(a, sep, flat, s);
    }
}

// Whether every pair's bin window over a 16 x 16 x kz tile fits a segment of seg elements: the
// window spans at most (farthest - nearest distance) * inv2 <= tile diagonal * inv2 bins, plus
// the 2-bin margins on each side and the extra element for the upper sample.
bool confocal_lds_fits(const float inc[9], float inv2, int kz, uint32_t seg) {
    double d2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        const double r = (kSegTileXY - 1.0) * (std::fabs((double)inc[k]) + std::fabs((double)inc[3 + k])) +
                         (kz - 1.0) * std::fabs((double)inc[6 + k]);
        d2 += r * r;
    }
    const double need = std::ceil(std::sqrt(d2) * std::fabs((double)inv2) * (1.0 + 1e-6)) + 8.0;
    return std::isfinite(need) && need <= (double)seg;
}
uint32_t lds_seg_elems(int) { return seg_elems(8); }   // every format stages 8-B elements
LdsShape lds_shape(int variant, int format) {
    // complex 16-bit rows accumulate two components per voxel; complex64 runs as two real passes
    const bool cplx = format == kCF16 || format == kCBF16;
    if (variant == kLdsVariant + 2) return {8, 8};                     // 512 lanes, 16^3 tiles
    if (cplx || variant != kLdsVariant) return {8, 4};
    return {16, 4};
}
bool confocal_lds_applies(const ConfocalArgs& a, int format, int variant, bool sep, bool flat) {
    return variant >= kLdsVariant && sep && flat &&
           confocal_lds_fits(a.g.inc, a.inv2, lds_shape(variant, format).depth(), lds_seg_elems(format));
}

template <class Fz, int KZ, int NW>
static hipError_t lds_launch(const ConfocalArgs& a, hipStream_t stream) {
    constexpr uint32_t TZ = KZ * NW / 4;
    const uint32_t n = ((a.g.n[0] + kSegTileXY - 1u) / kSegTileXY) * ((a.g.n[1] + kSegTileXY - 1u) / kSegTileXY) *
                       ((a.g.n[2] + TZ - 1u) / TZ);
    hipLaunchKernelGGL((confocal_lds_kernel<Fz, KZ, NW>), dim3(n), dim3(NW * 64), 0, stream, a);
    return hipGetLastError();
}
template <class Fz>
static hipError_t lds_launch_shape(const ConfocalArgs& a, LdsShape sh, hipStream_t stream) {
    if (sh.kz == 16) return lds_launch<Fz, 16, 4>(a, stream);
    if (sh.nw == 8) return lds_launch<Fz, 8, 8>(a, stream);
    return lds_launch<Fz, 8, 4>(a, stream);
}

hipError_t launch_confocal(const ConfocalArgs& a, int format, int variant, bool sep, bool flat,
                           hipStream_t stream) {
    if (confocal_lds_applies(a, format, variant, sep, flat)) {
        const LdsShape sh = lds_shape(variant, format);
        switch (format) {
            case kF32: return lds_launch_shape<FusedF32>(a, sh, stream);
            case kF16: return lds_launch_shape<FusedH16<false>>(a, sh, stream);
            case kBF16: return lds_launch_shape<FusedH16<true>>(a, sh, stream);
            case kC64: {   // two real passes, one per component, at the real formats' shapes
                const hipError_t e = lds_launch_shape<FusedC64Part<0>>(a, sh, stream);
                return e != hipSuccess ? e : lds_launch_shape<FusedC64Part<1>>(a, sh, stream);
            }
            case kCF16: return lds_launch<FusedC16<false>, 8, 4>(a, stream);
            case kCBF16: return lds_launch<FusedC16<true>, 8, 4>(a, stream);
            default: return hipErrorInvalidValue;
        }
    }
    if (variant_fused(variant)) {
        switch (format) {
            case kF32: return confocal_dispatch<FusedSource<FusedF32>>(a, variant, sep, flat, stream);
            case kF16: return confocal_dispatch<FusedSource<FusedH16<false>>>(a, variant, sep, flat, stream);
            case kBF16: return confocal_dispatch<FusedSource<FusedH16<true>>>(a, variant, sep, flat, stream);
            case kC64: return confocal_dispatch<FusedSource<FusedC64>>(a, variant, sep, flat, stream);
            case kCF16: return confocal_dispatch<FusedSource<FusedC16<false>>>(a, variant, sep, flat, stream);
            case kCBF16: return confocal_dispatch<FusedSource<FusedC16<true>>>(a, variant, sep, flat, stream);
            default: return hipErrorInvalidValue;
        }
    }
    switch (format) {
        case kF32: return confocal_dispatch<SplitSource<SmpF32>>(a, variant, sep, flat, stream);
        case kF16: return confocal_dispatch<SplitSource<SmpF16>>(a, variant, sep, flat, stream);
        case kBF16: return confocal_dispatch<SplitSource<SmpBF16>>(a, variant, sep, flat, stream);
        case kC64: return confocal_dispatch<SplitSource<SmpC64>>(a, variant, sep, flat, stream);
        case kCF16: return confocal_dispatch<SplitSource<SmpCF16>>(a, variant, sep, flat, stream);
        case kCBF16: return confocal_dispatch<SplitSource<SmpCBF16>>(a, variant, sep, flat, stream);
        default: return hipErrorInvalidValue;
    }
}

template <class Smp>
static hipError_t grid_dispatch(const GridArgs& a, int variant, hipStream_t s) {
    const size_t lds = sizeof(float) * (a.Tc + 1u);
    if (variant_k(variant) == 2)
        hipLaunchKernelGGL((grid_kernel<Smp, 2, 8>), dim3(a.t.count), dim3(256), lds, s, a);
    else
        hipLaunchKernelGGL((grid_kernel<Smp, 4, 8>), dim3(a.t.count), dim3(256), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_grid(const GridArgs& a, int format, int variant, hipStream_t stream) {
    switch (format) {
        case kF32: return grid_dispatch<SmpF32>(a, variant, stream);
        case kF16: return grid_dispatch<SmpF16>(a, variant, stream);
        case kBF16: return grid_dispatch<SmpBF16>(a, variant, stream);
        case kC64: return grid_dispatch<SmpC64>(a, variant, stream);
        case kCF16: return grid_dispatch<SmpCF16>(a, variant, stream);
        case kCBF16: return grid_dispatch<SmpCBF16>(a, variant, stream);
        default: return hipErrorInvalidValue;
    }
}

template <class Smp>
static hipError_t pairs_dispatch(const PairsArgs& a, hipStream_t s) {
    const size_t lds = sizeof(float) * (a.Tc + 1u);
    hipLaunchKernelGGL((pairs_kernel<Smp, 4>), dim3(a.t.count), dim3(256), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_pairs_raw(const PairsArgs& a, hipStream_t s) {
    const size_t lds = sizeof(float) * (a.Tc + 1u);
    hipLaunchKernelGGL((pairs_kernel<SmpF32, 4, true>), dim3(a.t.count), dim3(256), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_pairs(const PairsArgs& a, int format, int variant, hipStream_t stream) {
    (void)variant;
    switch (format) {
        case kF32: return pairs_dispatch<SmpF32>(a, stream);
        case kF16: return pairs_dispatch<SmpF16>(a, stream);
        case kBF16: return pairs_dispatch<SmpBF16>(a, stream);
        case kC64: return pairs_dispatch<SmpC64>(a, stream);
        case kCF16: return pairs_dispatch<SmpCF16>(a, stream);
        case kCBF16: return pairs_dispatch<SmpCBF16>(a, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace bpx
