// Plans: validation, structure detection, device staging and execution of one capture.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <thread>

#include "bpx_host.hpp"

namespace bpx {

// ------------------------------------------------------------------ device buffers
void DeviceBuffer::alloc(size_t n) {
    release();
    if (n == 0) n = 256;
    const hipError_t e = hipMalloc(&ptr, n);
    if (e != hipSuccess) {
        ptr = nullptr;
        fail(BPX_ERR_OUT_OF_MEMORY, "hipMalloc(" + std::to_string(n) + " bytes) failed: " +
                                        hipGetErrorString(e));
    }
    bytes = n;
}
void DeviceBuffer::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

PlanImpl::~PlanImpl() {
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    if (done) (void)hipEventDestroy(done);
    if (ready) (void)hipEventDestroy(ready);
    if (stream) (void)hipStreamDestroy(stream);
}

static uint32_t round_up(uint64_t x, uint32_t m) { return (uint32_t)((x + m - 1) / m * m); }

static bool finite_pair(const bpx_pair& p) {
    for (int k = 0; k < 3; ++k)
        if (!std::isfinite(p.camera_point[k]) || !std::isfinite(p.laser_point[k])) return false;
    return std::isfinite(p.camera_wall) && std::isfinite(p.laser_wall);
}

static bool same3(const float* a, const float* b) { return std::memcmp(a, b, 12) == 0; }

// Bound on |approximate quotient - exact quotient| of the table lookup.  The approximation
// q1 = fma(d', 2/dT, c1) (confocal) or fma(t, 1/dT, c1) differs from the exact fp32 chain
// (<= 3 rounded adds, 1 sqrt of <= 2 ulp error in the hardware sqrt, 1 rounded subtract,
// 1 rounded division) by a few units of 2^-24 times the magnitudes involved; 16 * 2^-24 * Q
// with Q = (max |total| + |t0|)/dT + Tc + 2 covers every term with a wide margin.
static double tap_error_bound(double max_total, double t0, double dT, double Tc) {
    const double Q = (std::fabs(max_total) + std::fabs(t0)) / dT + Tc + 2.0;
    return 16.0 * std::ldexp(1.0, -24) * Q;
}

static double dist3(const float* a, const double* b) {
    const double x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
    return std::sqrt(x * x + y * y + z * z);
}

// Voxel-grid corners of the A.1 grid (convex hull of all voxel positions).
static void grid_corners(const uint32_t n[3], const float zero[3], const float inc[9],
                         double corners[8][3]) {
    for (int m = 0; m < 8; ++m) {
        const double i[3] = {(m & 4) ? (double)(n[0] - 1) : 0.0, (m & 2) ? (double)(n[1] - 1) : 0.0,
                             (m & 1) ? (double)(n[2] - 1) : 0.0};
        for (int k = 0; k < 3; ++k)
            corners[m][k] = (double)zero[k] + inc[k] * i[0] + inc[3 + k] * i[1] + inc[6 + k] * i[2];
    }
}

// Per-pair element windows [lo, hi) of a fused confocal row (lde elements): every element a
// gather kernel can read for the pair and any voxel of the grid.  A voxel's distance d to the
// wall point lies between the distances to the nearest and farthest points of the corners'
// bounding box, the box widened by 1e-5 of the largest coordinate and the distances by a
// relative 1e-5 (the kernels' fp32 coordinates and squared distances are a few ulps off the
// exact ones); its bin is round((lw + 2 d + cw - t0) / deltaT) over that range, and the
// kernels read element bin + 1 and its neighbours (the two candidate bins of the approximate
// quotient, the second sample of a 16-bit element), so 4 elements of slack below and 5 above
// cover every read.  Bins before the row or past its end clamp to its ends (elements 0 and
// Tc, Tc + 1), so a window reaching past an end keeps that end's first 6 / last 6 elements,
// also when the whole window lies beyond it.
static std::vector<uint32_t> row_windows(const std::vector<bpx_pair>& pr, const double corners[8][3], double t0,
                                         double dT, uint32_t Tc, uint32_t lde) {
    double lo[3], hi[3], scale = 1.0;
    for (int k = 0; k < 3; ++k) lo[k] = hi[k] = corners[0][k];
    for (int m = 0; m < 8; ++m)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], corners[m][k]);
            hi[k] = std::max(hi[k], corners[m][k]);
            scale = std::max(scale, std::fabs(corners[m][k]));
        }
    for (const auto& p : pr)
        for (int k = 0; k < 3; ++k) scale = std::max(scale, std::fabs((double)p.laser_point[k]));
    for (int k = 0; k < 3; ++k) {
        lo[k] -= 1e-5 * scale;
        hi[k] += 1e-5 * scale;
    }
    std::vector<uint32_t> w(2 * pr.size());
    for (size_t p = 0; p < pr.size(); ++p) {
        double near2 = 0.0, far2 = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double x = pr[p].laser_point[k];
            const double c = x < lo[k] ? lo[k] - x : x > hi[k] ? x - hi[k] : 0.0;
            const double f = std::max(std::fabs(x - lo[k]), std::fabs(x - hi[k]));
            near2 += c * c;
            far2 += f * f;
        }
        const double base = (double)pr[p].laser_wall + (double)pr[p].camera_wall - t0;
        const double qlo = (base + 2.0 * std::sqrt(near2) * (1.0 - 1e-5)) / dT;
        const double qhi = (base + 2.0 * std::sqrt(far2) * (1.0 + 1e-5)) / dT;
        double elo = std::floor(qlo) + 1.0 - 4.0, ehi = std::ceil(qhi) + 1.0 + 5.0;
        if (!std::isfinite(elo) || !std::isfinite(ehi)) elo = 0.0, ehi = (double)lde;
        w[2 * p] = (uint32_t)std::clamp(elo, 0.0, (double)(Tc >= 4u ? Tc - 4u : 0u));
        w[2 * p + 1] = (uint32_t)std::clamp(ehi, (double)std::min(lde, 6u), (double)lde);
    }
    return w;
}

std::vector<uint32_t> confocal_row_windows(const bpx_pair* pairs, uint32_t P, const uint32_t n[3],
                                           const float zero[3], const float inc[9], float t0, float dT,
                                           uint32_t Tc) {
    double corners[8][3];
    grid_corners(n, zero, inc, corners);
    return row_windows(std::vector<bpx_pair>(pairs, pairs + P), corners, t0, dT, Tc, round_up((uint64_t)Tc + 2, 16));
}

static double max_total_path(const bpx_pair* pr, size_t count, const double corners[8][3]) {
    double mx = 0.0;
    for (size_t i = 0; i < count; ++i) {
        const bpx_pair& p = pr[i];
        double dl = 0, dc = 0;
        for (int m = 0; m < 8; ++m) {
            dl = std::max(dl, dist3(p.laser_point, corners[m]));
            dc = std::max(dc, dist3(p.camera_point, corners[m]));
        }
        mx = std::max(mx, std::fabs((double)p.laser_wall) + std::fabs((double)p.camera_wall) + dl + dc);
    }
    return mx;
}

double single_tap_error_bound(const bpx_pair* pairs, uint32_t P, const uint32_t n[3], const float zero[3],
                              const float inc[9], float t0, float dT, uint32_t T) {
    double corners[8][3];
    grid_corners(n, zero, inc, corners);
    return tap_error_bound(max_total_path(pairs, P, corners), t0, dT, T);
}

// Power-of-two scale s so that max|x| / s lies in [2^14, 2^15) (exact rescaling).
// Largest |sample| of the chunk (host chunk, or the copied windows of the source rows).
static float chunk_max_abs(const CaptureIn& c) {
    float m = 0.0f;
    if (c.chunk) {
        const uint64_t n = c.size * (c.cplx ? 2 : 1);
        for (uint64_t i = 0; i < n; ++i) m = std::max(m, std::fabs(c.chunk[i]));
    } else if (c.device_rows) {
        const uint64_t n = c.size * (c.cplx ? 2 : 1);
        DeviceBuffer r;
        r.alloc(sizeof(float));
        hip_check(hipMemset(r.ptr, 0, sizeof(float)), "hipMemset");
        hip_check(launch_abs_max(static_cast<const float*>(c.device_rows->ptr), n, static_cast<float*>(r.ptr),
                                 nullptr),
                  "abs max kernel");
        hip_check(hipMemcpy(&m, r.ptr, sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
    } else if (c.src) {
        for (uint32_t p = 0; p < c.P; ++p) {
            const float* r = c.src + (size_t)p * c.src_T + c.from;
            for (uint32_t i = 0; i < c.ncopy; ++i) m = std::max(m, std::fabs(r[i]));
        }
    }
    return m;
}

static float choose_scale(const CaptureIn& c) {
    const float m = chunk_max_abs(c);
    if (!(m > 0.0f) || !std::isfinite(m)) return 1.0f;
    int e;
    std::frexp(m, &e);  // m in [2^(e-1), 2^e)
    return std::ldexp(1.0f, e - 1 - 14);
}

static void validate_common(const CaptureIn& c, const uint32_t n[3]) {
    if (!c.chunk && !c.src && !c.device_rows && c.size) fail(BPX_ERR_INVALID_ARGUMENT, "transient_chunk is NULL");
    if (c.device_rows && c.device_rows->bytes < sizeof(float) * c.size * (c.cplx ? 2 : 1))
        fail(BPX_ERR_INVALID_ARGUMENT, "device rows smaller than the chunk");
    if (!c.chunk && c.src) {
        if (c.cplx) fail(BPX_ERR_INVALID_ARGUMENT, "device cropping takes real samples");
        if ((uint64_t)c.from + c.ncopy > c.src_T || (uint64_t)c.cfrom + c.ncopy > c.T)
            fail(BPX_ERR_INVALID_ARGUMENT, "crop window outside the source or chunk rows");
    }
    if (!c.pairs && c.P) fail(BPX_ERR_INVALID_ARGUMENT, "scanned_pairs is NULL");
    if (c.T == 0) fail(BPX_ERR_INVALID_ARGUMENT, "T must be > 0");
    if ((uint64_t)c.P * c.T != c.size)
        fail(BPX_ERR_INVALID_ARGUMENT, "transient_size (" + std::to_string(c.size) +
                                           ") != num_pairs * T (" +
                                           std::to_string((uint64_t)c.P * c.T) + ")");
    if (!std::isfinite(c.t0)) fail(BPX_ERR_INVALID_ARGUMENT, "t0 must be finite");
    if (!(c.dT > 0.0f) || !std::isfinite(c.dT) || !std::isfinite(1.0f / c.dT) ||
        c.dT < std::numeric_limits<float>::min())
        fail(BPX_ERR_INVALID_ARGUMENT, "deltaT must be a positive normal float");
    for (int k = 0; k < 3; ++k)
        if (n[k] == 0) fail(BPX_ERR_INVALID_ARGUMENT, "voxels_per_side must be > 0");
    if ((uint64_t)n[0] * n[1] * n[2] > (1ull << 32))
        fail(BPX_ERR_UNSUPPORTED, "volumes above 2^32 voxels are not supported");
    if (c.T > (1u << 22)) fail(BPX_ERR_UNSUPPORTED, "T above 2^22 bins is not supported");
}

// Pairs that take part: finite ones, minus the reference launcher's dropped tail if asked.
static std::vector<uint32_t> used_pairs(const CaptureIn& c, uint32_t flags) {
    uint32_t P = c.P;
    if (flags & BPX_FLAG_REFERENCE_PAIR_TRUNCATION) {
        const uint32_t B = std::min<uint32_t>(256u, std::max<uint32_t>(P, 1u));
        P = (P / B) * B;
    }
    std::vector<uint32_t> idx;
    idx.reserve(P);
    for (uint32_t p = 0; p < P; ++p)
        if (finite_pair(c.pairs[p])) idx.push_back(p);
    return idx;
}

// ------------------------------------------------------------------ host -> device uploads
// Large pageable host arrays go through two pinned staging buffers (allocated once per
// process): host threads copy chunk k + 1 into one buffer while the DMA engine copies chunk k
// from the other.  Measured on MI355X: pageable hipMemcpy ~30 GB/s, pinned ~57 GB/s, 4-thread
// host copies into pinned memory ~60 GB/s.
namespace {
constexpr size_t kStageChunk = 32u << 20;
constexpr int kCopyThreads = 4;
std::mutex g_upload_mutex;
void* g_pinned[2] = {nullptr, nullptr};
hipEvent_t g_pinned_ev[2] = {nullptr, nullptr};

void parallel_copy(void* dst, const void* src, size_t bytes) {
    std::thread th[kCopyThreads - 1];
    const size_t part = (bytes / kCopyThreads + 63) & ~size_t(63);
    auto piece = [&](int t) {
        const size_t a = std::min(bytes, part * (size_t)t), b = std::min(bytes, part * (size_t)(t + 1));
        if (b > a) std::memcpy(static_cast<char*>(dst) + a, static_cast<const char*>(src) + a, b - a);
    };
    for (int t = 1; t < kCopyThreads; ++t) th[t - 1] = std::thread(piece, t);
    piece(0);
    for (auto& x : th) x.join();
}
}  // namespace

void upload_host(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes < 2 * kStageChunk) {
        hip_check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy(upload)");
        return;
    }
    std::lock_guard<std::mutex> lock(g_upload_mutex);
    for (int b = 0; b < 2; ++b) {
        if (!g_pinned[b]) hip_check(hipHostMalloc(&g_pinned[b], kStageChunk, hipHostMallocDefault), "hipHostMalloc");
        if (!g_pinned_ev[b]) hip_check(hipEventCreateWithFlags(&g_pinned_ev[b], hipEventDisableTiming), "hipEventCreate");
    }
    size_t off = 0;
    for (int k = 0; off < bytes; ++k, off += kStageChunk) {
        const int b = k & 1;
        const size_t n = std::min(kStageChunk, bytes - off);
        hip_check(hipEventSynchronize(g_pinned_ev[b]), "hipEventSynchronize");   // its previous DMA is done
        parallel_copy(g_pinned[b], static_cast<const char*>(src) + off, n);
        hip_check(hipMemcpyAsync(static_cast<char*>(dst) + off, g_pinned[b], n, hipMemcpyHostToDevice, s),
                  "hipMemcpyAsync(upload)");
        hip_check(hipEventRecord(g_pinned_ev[b], s), "hipEventRecord");
    }
    // the buffers may be reused as soon as this returns
    hip_check(hipEventSynchronize(g_pinned_ev[0]), "hipEventSynchronize");
    hip_check(hipEventSynchronize(g_pinned_ev[1]), "hipEventSynchronize");
}

// Stages the raw rows of the used pairs on the device (compacting when pairs were dropped).
// Device rows [rows_in][row floats] -> the plan's raw rows of the used pairs: taken over when
// every row is used, else gathered.
static void adopt_device_rows(PlanImpl& pl, DeviceBuffer& rows, uint32_t rows_in, size_t row,
                              const std::vector<uint32_t>& idx) {
    if (idx.size() == rows_in && rows_in) {
        pl.raw.release();
        std::swap(pl.raw.ptr, rows.ptr);
        std::swap(pl.raw.bytes, rows.bytes);
        return;
    }
    pl.raw.alloc(sizeof(float) * row * std::max<size_t>(idx.size(), 1));
    if (idx.empty()) return;
    DeviceBuffer map;
    map.alloc(sizeof(uint32_t) * idx.size());
    hip_check(hipMemcpy(map.ptr, idx.data(), sizeof(uint32_t) * idx.size(), hipMemcpyHostToDevice),
              "hipMemcpy(row map)");
    hip_check(launch_crop_rows(static_cast<const float*>(rows.ptr), (uint32_t)row, static_cast<const uint32_t*>(map.ptr),
                               (uint32_t)idx.size(), (uint32_t)row, 0u, 0u, (uint32_t)row,
                               static_cast<float*>(pl.raw.ptr), pl.stream),
              "row gather kernel");
    hip_check(hipStreamSynchronize(pl.stream), "hipStreamSynchronize");
}

void plan_replace_rows(PlanImpl& pl, DeviceBuffer& rows) {
    if (pl.path == kPathOctree) fail(BPX_ERR_UNSUPPORTED, "octree plans keep their rows");
    if (pl.P == 0 || !pl.raw.ptr) return;   // nothing staged: every execution writes zeros
    const size_t row = (size_t)pl.T * (pl.cplx ? 2 : 1);
    if (rows.bytes < sizeof(float) * row * pl.rows_in) fail(BPX_ERR_INVALID_ARGUMENT, "replacement rows too small");
    plan_wait(pl);
    adopt_device_rows(pl, rows, pl.rows_in, row, pl.used);
    if (pl.stage.ptr) {   // the staging descriptor names the raw rows: point it at the new ones
        const StageJob job = stage_job(pl);
        hip_check(hipMemcpy(pl.stage.ptr, &job, sizeof(StageJob), hipMemcpyHostToDevice), "hipMemcpy(staging job)");
    }
}

static void stage_raw(PlanImpl& pl, const CaptureIn& c, const std::vector<uint32_t>& idx) {
    const size_t comp = c.cplx ? 2 : 1;
    const size_t row = (size_t)c.T * comp;
    pl.used = idx;
    pl.rows_in = c.P;
    if (!c.chunk && c.device_rows) {
        adopt_device_rows(pl, *c.device_rows, c.P, row, idx);
        return;
    }
    pl.raw.alloc(sizeof(float) * row * std::max<size_t>(idx.size(), 1));
    if (!c.chunk) {
        if (idx.empty()) return;
        if (idx.size() == c.P && c.from == 0 && c.cfrom == 0 && c.ncopy == c.T && c.src_T == c.T) {
            upload_host(pl.raw.ptr, c.src, sizeof(float) * row * c.P, pl.stream);
            return;
        }
        // the source rows in one copy, then the crop (and compaction of dropped pairs) on the device
        DeviceBuffer full, map;
        full.alloc(sizeof(float) * (size_t)c.P * c.src_T);
        map.alloc(sizeof(uint32_t) * idx.size());
        upload_host(full.ptr, c.src, sizeof(float) * (size_t)c.P * c.src_T, pl.stream);
        hip_check(hipMemcpy(map.ptr, idx.data(), sizeof(uint32_t) * idx.size(), hipMemcpyHostToDevice),
                  "hipMemcpy(row map)");
        hip_check(launch_crop_rows(static_cast<const float*>(full.ptr), c.src_T,
                                   static_cast<const uint32_t*>(map.ptr), (uint32_t)idx.size(), c.T, c.from,
                                   c.cfrom, c.ncopy, static_cast<float*>(pl.raw.ptr), pl.stream),
                  "crop kernel");
        hip_check(hipStreamSynchronize(pl.stream), "hipStreamSynchronize");
        return;
    }
    if (idx.size() == c.P) {
        upload_host(pl.raw.ptr, c.chunk, sizeof(float) * row * c.P, pl.stream);
    } else {
        std::vector<float> tmp(row * idx.size());
        for (size_t i = 0; i < idx.size(); ++i)
            std::memcpy(&tmp[i * row], c.chunk + (size_t)idx[i] * row, sizeof(float) * row);
        if (!tmp.empty())
            hip_check(hipMemcpy(pl.raw.ptr, tmp.data(), sizeof(float) * tmp.size(),
                                hipMemcpyHostToDevice),
                      "hipMemcpy(transient)");
    }
}

static void make_events(PlanImpl& pl) {
    for (auto& e : pl.ev) hip_check(hipEventCreate(&e), "hipEventCreate");
    hip_check(hipEventCreateWithFlags(&pl.done, hipEventDisableTiming), "hipEventCreate");
    hip_check(hipEventCreateWithFlags(&pl.ready, hipEventDisableTiming), "hipEventCreate");
}

void plan_wait(const PlanImpl& pl) {
    hip_check(hipStreamSynchronize(pl.stream), "hipStreamSynchronize");
    if (pl.done) hip_check(hipEventSynchronize(pl.done), "hipEventSynchronize");
}

// Exhaustive-grid structure: p = l*C + c with laser data depending only on l and camera
// data only on c.  Returns C (0 when the list is not such a product).
static uint32_t grid_structure(const std::vector<bpx_pair>& pr) {
    const uint32_t P = (uint32_t)pr.size();
    if (P < 2) return 0;
    uint32_t C = 1;
    while (C < P && same3(pr[C].laser_point, pr[0].laser_point) &&
           pr[C].laser_wall == pr[0].laser_wall)
        ++C;
    if (C == P || P % C != 0 || C < 2) return 0;
    for (uint32_t p = 0; p < P; ++p) {
        const uint32_t l = p / C, c = p % C;
        const bpx_pair& a = pr[p];
        const bpx_pair& lb = pr[(size_t)l * C];
        const bpx_pair& cb = pr[c];
        if (!same3(a.laser_point, lb.laser_point) || a.laser_wall != lb.laser_wall) return 0;
        if (!same3(a.camera_point, cb.camera_point) || a.camera_wall != cb.camera_wall) return 0;
    }
    return C;
}

PlanImpl* make_plan_single_tap(const CaptureIn& c, const uint32_t n[3], const float zero[3],
                               const float inc[9], const bpx_options* opt) {
    validate_common(c, n);
    for (int k = 0; k < 3; ++k)
        if (!std::isfinite(zero[k])) fail(BPX_ERR_INVALID_ARGUMENT, "volume_zero_pos must be finite");
    for (int k = 0; k < 9; ++k)
        if (!std::isfinite(inc[k])) fail(BPX_ERR_INVALID_ARGUMENT, "voxel_inc must be finite");
    if (!device_ready()) fail(BPX_ERR_NO_DEVICE, "no gfx950 device available");

    bpx_options o{BPX_STORAGE_F32, 32u, 0u, -1};
    if (opt) o = *opt;
    if (o.kahan_block == 0) o.kahan_block = 32;
    if (o.storage > BPX_STORAGE_BF16) fail(BPX_ERR_INVALID_ARGUMENT, "unknown storage format");
    if (o.variant >= variant_count()) fail(BPX_ERR_INVALID_ARGUMENT, "unknown kernel variant");

    auto pl = std::make_unique<PlanImpl>();
    hip_check(hipStreamCreateWithFlags(&pl->stream, hipStreamNonBlocking), "hipStreamCreate");
    make_events(*pl);
    pl->cplx = c.cplx;
    pl->format = c.cplx ? (o.storage == BPX_STORAGE_F16 ? kCF16 : o.storage == BPX_STORAGE_BF16 ? kCBF16 : kC64)
                        : (o.storage == BPX_STORAGE_F16 ? kF16 : o.storage == BPX_STORAGE_BF16 ? kBF16 : kF32);
    pl->kahan_block = o.kahan_block;
    pl->t0 = c.t0;
    pl->dT = c.dT;
    pl->T = c.T;
    std::memcpy(pl->g.n, n, sizeof(uint32_t) * 3);
    std::memcpy(pl->g.zero, zero, sizeof(float) * 3);
    std::memcpy(pl->g.inc, inc, sizeof(float) * 9);

    const std::vector<uint32_t> idx = used_pairs(c, o.flags);
    std::vector<bpx_pair> pr(idx.size());
    for (size_t i = 0; i < idx.size(); ++i) pr[i] = c.pairs[idx[i]];
    pl->P = (uint32_t)pr.size();

    const size_t nvox = (size_t)n[0] * n[1] * n[2];
    pl->vol.alloc(nvox * (c.cplx ? 8 : 4));

    // structure
    bool confocal = !pr.empty();
    for (const auto& p : pr) confocal = confocal && same3(p.laser_point, p.camera_point);
    pl->C = confocal ? 0 : grid_structure(pr);
    pl->L = pl->C ? pl->P / pl->C : 0;
    pl->path = confocal ? kPathConfocal : (pl->C ? kPathGrid : kPathPairs);

    // exactness precondition of the table lookup
    double corners[8][3];
    grid_corners(n, zero, inc, corners);
    const double bound = tap_error_bound(max_total_path(pr.data(), pr.size(), corners), c.t0, c.dT, c.T);
    if (bound >= 0.25)
        fail(BPX_ERR_UNSUPPORTED,
             "path lengths / t0 too large relative to deltaT for the exact bin lookup (bound " +
                 std::to_string(bound) + ")");

    pl->variant = o.variant < 0 ? kDefaultVariant : o.variant;
    pl->fused = pl->path == kPathConfocal && variant_fused(pl->variant);
    pl->K = variant_k(pl->variant);
    if (pl->path == kPathPairs) pl->K = 4;
    if (pl->path == kPathGrid && pl->K == 8) pl->K = 4;
    pl->tiles = make_tiles(n, pl->K);

    if (pl->P == 0) return pl.release();

    // samples: raw rows resident on the device; scale for 16-bit storage
    pl->scale = (pl->format == kF32 || pl->format == kC64) ? 1.0f : choose_scale(c);
    stage_raw(*pl, c, idx);
    // padded rows (split) or the fused sample+threshold layout, [P][ldh] elements
    pl->ldh = round_up((uint64_t)c.T + 2, pl->fused ? 16 : 64);
    const uint64_t hbytes = (uint64_t)pl->P * pl->ldh *
                            (pl->fused ? fused_elem_bytes(pl->format) : format_bytes(pl->format));
    if (hbytes >= (1ull << 32)) fail(BPX_ERR_UNSUPPORTED, "transient rows above 4 GiB per plan");
    pl->H.alloc(hbytes);
    // fused rows staged over per-pair windows keep zeros outside them (written here only)
    if (pl->fused) hip_check(hipMemset(pl->H.ptr, 0, hbytes), "hipMemset(fused rows)");

    std::vector<float> p8((size_t)pl->P * 8);
    std::memcpy(p8.data(), pr.data(), sizeof(float) * p8.size());
    pl->pair8.alloc(sizeof(float) * p8.size());
    hip_check(hipMemcpy(pl->pair8.ptr, p8.data(), sizeof(float) * p8.size(), hipMemcpyHostToDevice),
              "hipMemcpy(pairs)");

    if (pl->path == kPathConfocal) {
        pl->Z.alloc(sizeof(float) * ((size_t)c.T + 1));
        if (!pl->fused) {
            pl->lds = round_up((uint64_t)c.T + 1, 64);
            const uint64_t sbytes = (uint64_t)pl->P * pl->lds * 4;
            if (sbytes >= (1ull << 32)) fail(BPX_ERR_UNSUPPORTED, "threshold table above 4 GiB");
            pl->S.alloc(sbytes);
        } else {
            pl->win_host = row_windows(pr, corners, c.t0, c.dT, c.T, pl->ldh);
            pl->win.alloc(sizeof(uint32_t) * pl->win_host.size());
            hip_check(hipMemcpy(pl->win.ptr, pl->win_host.data(), sizeof(uint32_t) * pl->win_host.size(),
                                hipMemcpyHostToDevice),
                      "hipMemcpy(row windows)");
            const StageJob job = stage_job(*pl);
            pl->stage_zb = stage_z_blocks(job);
            pl->stage_sb = stage_row_blocks(job);
            pl->stage.alloc(sizeof(StageJob));
            hip_check(hipMemcpy(pl->stage.ptr, &job, sizeof(StageJob), hipMemcpyHostToDevice),
                      "hipMemcpy(staging job)");
        }
        pl->inv2 = (float)(2.0 / (double)c.dT);
        std::vector<float> p4((size_t)pl->P * 4);
        bool flat = true;
        for (uint32_t p = 0; p < pl->P; ++p) {
            p4[4 * p + 0] = pr[p].laser_point[0];
            p4[4 * p + 1] = pr[p].laser_point[1];
            p4[4 * p + 2] = pr[p].laser_point[2];
            p4[4 * p + 3] = (float)(((double)pr[p].laser_wall + (double)pr[p].camera_wall - (double)c.t0) /
                                        (double)c.dT + 1.0);
            flat = flat && pr[p].laser_point[2] == pr[0].laser_point[2];
        }
        pl->flat = flat;
        pl->zflat = flat;
        pl->lz = pr[0].laser_point[2];
        pl->sep = inc[6] == 0.0f && inc[7] == 0.0f;
        if (!pl->sep) pl->flat = false;
        // Default kernel: the LDS-staged one where it applies (separable grid, flat wall, short
        // bin windows) and has enough tiles to fill the GPU.  Real formats: 16^3 tiles of 256
        // lanes (16 voxels per lane) from 1,024 tiles, 16^3 tiles of 512 lanes (8 per lane)
        // from 512; else, and for complex formats, 16 x 16 x 8 tiles of 256 lanes from 512.
        // Measured at 16,384 pairs (gather ms, the three shapes in that order, texture path
        // last): 256^3 81.7 / 85.8 / 96.7 / 141; 128^3 16.1 / 12.8 / 14.4 / 20.0; 112^3 - /
        // 13.2 / 12.1 / 19.4; 96^3 - / - / 10.2 / 10.0.
        const uint32_t seg = lds_seg_elems(pl->format);
        if (o.variant < 0 && pl->fused && pl->sep && pl->flat) {
            const uint64_t t16 = (uint64_t)((n[0] + 15) / 16) * ((n[1] + 15) / 16) * ((n[2] + 15) / 16);
            const uint64_t t8 = (uint64_t)((n[0] + 15) / 16) * ((n[1] + 15) / 16) * ((n[2] + 7) / 8);
            const bool cplx = pl->format == kCF16 || pl->format == kCBF16;   // two accumulators per voxel
            const bool fits16 = confocal_lds_fits(inc, pl->inv2, 16, seg), fits8 = confocal_lds_fits(inc, pl->inv2, 8, seg);
            if (!cplx && t16 >= kLdsDeepTiles && fits16) pl->variant = kLdsVariant;
            else if (!cplx && t16 >= kLdsMinTiles && fits16) pl->variant = kLdsVariant + 2;
            else if (t8 >= kLdsMinTiles && fits8) pl->variant = kLdsVariant + 1;
        }
        pl->keep_f32 = (o.flags & BPX_FLAG_KEEP_F32_ROWS) != 0;
        if (pl->fused) {
            set_fused_variant(*pl, pl->variant);
        } else {
            pl->lds_tiles = false;
            pl->rows16 = false;
        }
        pl->pair4.alloc(sizeof(float) * p4.size());
        hip_check(hipMemcpy(pl->pair4.ptr, p4.data(), sizeof(float) * p4.size(), hipMemcpyHostToDevice),
                  "hipMemcpy(pair table)");
    } else {
        if ((uint64_t)(c.T + 1) * 4 > 64 * 1024)
            fail(BPX_ERR_UNSUPPORTED, "non-confocal captures need T <= 16383 after cropping");
        pl->Z.alloc(sizeof(float) * (c.T + 1));
        pl->inv = (float)(1.0 / (double)c.dT);
        pl->c1 = (float)(1.0 - (double)c.t0 / (double)c.dT);
        if (pl->path == kPathGrid) {
            std::vector<float> las((size_t)pl->L * 4), spd((size_t)pl->C * 4);
            for (uint32_t l = 0; l < pl->L; ++l) {
                const bpx_pair& a = pr[(size_t)l * pl->C];
                las[4 * l + 0] = a.laser_point[0];
                las[4 * l + 1] = a.laser_point[1];
                las[4 * l + 2] = a.laser_point[2];
                las[4 * l + 3] = a.laser_wall;
            }
            for (uint32_t s = 0; s < pl->C; ++s) {
                spd[4 * s + 0] = pr[s].camera_point[0];
                spd[4 * s + 1] = pr[s].camera_point[1];
                spd[4 * s + 2] = pr[s].camera_point[2];
                spd[4 * s + 3] = pr[s].camera_wall;
            }
            pl->lasers.alloc(sizeof(float) * las.size());
            pl->spads.alloc(sizeof(float) * spd.size());
            hip_check(hipMemcpy(pl->lasers.ptr, las.data(), sizeof(float) * las.size(), hipMemcpyHostToDevice),
                      "hipMemcpy(lasers)");
            hip_check(hipMemcpy(pl->spads.ptr, spd.data(), sizeof(float) * spd.size(), hipMemcpyHostToDevice),
                      "hipMemcpy(spads)");
        }
    }
    return pl.release();
}

int staged_format(const PlanImpl& pl) {
    if (!pl.rows16) return pl.format;
    plan_wait(pl);
    uint32_t flag = 1;
    hip_check(hipMemcpy(&flag, pl.fmt_flag.ptr, 4, hipMemcpyDeviceToHost), "hipMemcpy(row format)");
    return flag == 0u ? kF16 : pl.format;
}

StageJob stage_job(const PlanImpl& pl) {
    StageJob j{};
    j.src = static_cast<const float*>(pl.raw.ptr);
    j.dst = pl.H.ptr;
    j.pair8 = static_cast<const float*>(pl.pair8.ptr);
    j.Z = static_cast<float*>(pl.Z.ptr);
    j.rows = pl.P;
    j.Tc = pl.T;
    j.lde = pl.ldh;
    j.format = pl.format;
    j.fmt_flag = pl.rows16 ? static_cast<const uint32_t*>(pl.fmt_flag.ptr) : nullptr;
    j.win = static_cast<const uint32_t*>(pl.win.ptr);
    j.alt_format = kF16;
    j.inv_scale = 1.0f / pl.scale;
    j.t0 = pl.t0;
    j.dT = pl.dT;
    return j;
}

void set_fused_variant(PlanImpl& pl, int variant) {
    pl.variant = variant;
    pl.K = variant_k(variant);
    pl.tiles = make_tiles(pl.g.n, pl.K);
    pl.lds_tiles = variant >= kLdsVariant && pl.sep && pl.flat &&
                   confocal_lds_fits(pl.g.inc, pl.inv2, lds_shape(variant, pl.format).depth(),
                                     lds_seg_elems(pl.format));
    // fp32 rows of integer samples within fp16's exact range are staged as fp16: the same
    // values, half the LDS bytes per lookup (one 8-B read instead of an 8-B and a 4-B one).
    // Each execution checks its samples on the device first and stages and gathers in the
    // format the check allows, so the volume is the fp32 rows' bit for bit either way.  For
    // the shapes of 8 voxels per lane (gather ms at 16,384 pairs, fp32 -> fp16 rows: 128^3
    // 16^3 x 512 lanes 12.96 -> 11.67, 16 x 16 x 8 x 256 14.57 -> 13.92; 256^3 83.3 -> 83.2,
    // 91.7 -> 89.7); the 16-per-lane shape, bound by its arithmetic and the default where
    // it applies (256^3), measured slower with fp16 rows (84.5 -> 86.4), so it keeps fp32.
    pl.rows16 = pl.lds_tiles && pl.format == kF32 && lds_shape(variant, pl.format).kz == 8 && !pl.keep_f32;
    if (pl.rows16 && !pl.fmt_flag.ptr) {
        pl.fmt_flag.alloc(4);
        hip_check(hipMemset(pl.fmt_flag.ptr, 0xff, 4), "hipMemset(row format)");   // fp32 until run
    }
    if (pl.stage.ptr) {   // the staging descriptor names the format flag (or none)
        const StageJob job = stage_job(pl);
        hip_check(hipMemcpy(pl.stage.ptr, &job, sizeof(StageJob), hipMemcpyHostToDevice), "hipMemcpy(staging job)");
    }
}

void stage_rows(PlanImpl& pl, hipStream_t s) {
    hip_check(launch_stage(static_cast<const StageJob*>(pl.stage.ptr), 1u, pl.stage_zb, pl.stage_sb, s),
              "row staging kernels");
}

void PlanImpl::execute(hipStream_t s) {
    if (path == kPathOctree) {
        execute_octree(*this, s);
        return;
    }
    const size_t nvox = (size_t)g.n[0] * g.n[1] * g.n[2];
    if (path == kPathWindow) {
        if (P == 0 || T < 2) {
            hip_check(hipMemsetAsync(vol.ptr, 0, nvox * (cplx ? 8 : 4), s), "hipMemsetAsync");
            return;
        }
        hip_check(hipEventRecord(ev[0], s), "hipEventRecord");
        hip_check(launch_prefix(static_cast<const float*>(raw.ptr), P, T, cplx ? 1 : 0, H.ptr, ldh, s),
                  "prefix kernel");
        hip_check(hipEventRecord(ev[1], s), "hipEventRecord");
        hip_check(launch_ztable(1, t0, dT, half_width, T, static_cast<float*>(Z.ptr), s), "window table (start)");
        hip_check(launch_ztable(2, t0, dT, half_width, T, static_cast<float*>(Z2.ptr), s), "window table (end)");
        hip_check(hipEventRecord(ev[2], s), "hipEventRecord");
        WindowArgs a{};
        a.PS = H.ptr;
        a.pairs = static_cast<const float*>(pair8.ptr);
        a.Zlo = static_cast<const float*>(Z.ptr);
        a.Zhi = static_cast<const float*>(Z2.ptr);
        a.out = vol.ptr;
        a.P = P;
        a.T = T;
        a.ldp = ldh;
        a.ps_bytes = H.bytes;
        std::memcpy(a.n, g.n, sizeof(a.n));
        std::memcpy(a.zero, wzero, sizeof(a.zero));
        std::memcpy(a.vsize, wvs, sizeof(a.vsize));
        a.inv = inv;
        a.c1_lo = c1_lo;
        a.c1_hi = c1_hi;
        a.confocal = all_confocal ? 1u : 0u;
        hip_check(launch_window(a, cplx ? 1 : 0, s), "window gather kernel");
        hip_check(hipEventRecord(ev[3], s), "hipEventRecord");
        return;
    }
    if (P == 0) {
        hip_check(hipMemsetAsync(vol.ptr, 0, nvox * (cplx ? 8 : 4), s), "hipMemsetAsync");
        return;
    }
    // events: ev[0] | split-layout row padding | ev[1] | tables (fused: all staging) | ev[2] |
    // gather | ev[3]
    hip_check(hipEventRecord(ev[0], s), "hipEventRecord");
    if (!fused)
        hip_check(launch_pad_rows(static_cast<const float*>(raw.ptr), P, T, T, H.ptr, ldh, format,
                                  1.0f / scale, s),
                  "row staging kernel");
    hip_check(hipEventRecord(ev[1], s), "hipEventRecord");
    if (path == kPathConfocal) {
        if (fused) {
            if (rows16) {
                hip_check(launch_fill(static_cast<float*>(fmt_flag.ptr), 1, 0.0f, s), "fill kernel");
                hip_check(launch_exact16_check(static_cast<const float*>(raw.ptr), (uint64_t)P * T,
                                               static_cast<uint32_t*>(fmt_flag.ptr), s),
                          "exact fp16 check kernel");
            }
            stage_rows(*this, s);
        } else {
            hip_check(launch_ztable(0, t0, dT, 0.0f, T + 1, static_cast<float*>(Z.ptr), s),
                      "path-length table kernel");
            hip_check(launch_stable(static_cast<const float*>(pair8.ptr), P, T, static_cast<const float*>(Z.ptr),
                                    static_cast<float*>(S.ptr), lds, 1u, 0u, s),
                      "threshold table kernel");
        }
        hip_check(hipEventRecord(ev[2], s), "hipEventRecord");
        ConfocalArgs a{};
        a.H = H.ptr;
        a.S = static_cast<const float*>(S.ptr);
        a.pairs = static_cast<const float4*>(pair4.ptr);
        a.out = vol.ptr;
        a.P = P;
        a.Tc = T;
        a.ldh = ldh;
        a.lds = lds;
        a.h_bytes = (uint32_t)H.bytes;
        a.s_bytes = (uint32_t)(fused ? 0 : S.bytes);
        a.g = g;
        a.t = tiles;
        a.inv2 = inv2;
        a.scale = scale;
        a.lz = lz;
        a.kahan_block = kahan_block;
        if (rows16) {
            // both instantiations launched; the one the device check did not pick returns at once
            a.fmt_flag = static_cast<const uint32_t*>(fmt_flag.ptr);
            a.fmt_want = 1u;
            hip_check(launch_confocal(a, format, variant, sep, flat, s), "confocal gather kernel");
            a.fmt_want = 0u;
            hip_check(launch_confocal(a, kF16, variant, sep, flat, s), "confocal gather kernel (fp16 rows)");
        } else {
            hip_check(launch_confocal(a, format, variant, sep, flat, s), "confocal gather kernel");
        }
    } else {
        hip_check(launch_ztable(0, t0, dT, 0.0f, T + 1, static_cast<float*>(Z.ptr), s),
                  "threshold table kernel");
        hip_check(hipEventRecord(ev[2], s), "hipEventRecord");
        if (path == kPathGrid) {
            GridArgs a{};
            a.H = H.ptr;
            a.lasers = static_cast<const float4*>(lasers.ptr);
            a.spads = static_cast<const float4*>(spads.ptr);
            a.Z = static_cast<const float*>(Z.ptr);
            a.out = vol.ptr;
            a.L = L;
            a.C = C;
            a.Tc = T;
            a.ldh = ldh;
            a.h_bytes = (uint32_t)H.bytes;
            a.g = g;
            a.t = tiles;
            a.inv = inv;
            a.c1 = c1;
            a.scale = scale;
            a.kahan_block = kahan_block;
            hip_check(launch_grid(a, format, variant, s), "grid gather kernel");
        } else {
            PairsArgs a{};
            a.H = H.ptr;
            a.pairs = static_cast<const float*>(pair8.ptr);
            a.Z = static_cast<const float*>(Z.ptr);
            a.out = vol.ptr;
            a.P = P;
            a.Tc = T;
            a.ldh = ldh;
            a.h_bytes = (uint32_t)H.bytes;
            a.g = g;
            a.t = tiles;
            a.inv = inv;
            a.c1 = c1;
            a.scale = scale;
            a.kahan_block = kahan_block;
            hip_check(launch_pairs(a, format, variant, s), "pairs gather kernel");
        }
    }
    hip_check(hipEventRecord(ev[3], s), "hipEventRecord");
}

PlanImpl* make_plan_window(const CaptureIn& c, const uint32_t n[3], const float vpos[3],
                           const float vsize[3]) {
    validate_common(c, n);
    for (int k = 0; k < 3; ++k)
        if (!std::isfinite(vpos[k]) || !std::isfinite(vsize[k]))
            fail(BPX_ERR_INVALID_ARGUMENT, "volume position/size must be finite");
    if (!device_ready()) fail(BPX_ERR_NO_DEVICE, "no gfx950 device available");
    auto pl = std::make_unique<PlanImpl>();
    hip_check(hipStreamCreateWithFlags(&pl->stream, hipStreamNonBlocking), "hipStreamCreate");
    make_events(*pl);
    pl->path = kPathWindow;
    pl->cplx = c.cplx;
    pl->format = c.cplx ? kC64 : kF32;
    pl->t0 = c.t0;
    pl->dT = c.dT;
    pl->T = c.T;
    std::memcpy(pl->g.n, n, sizeof(uint32_t) * 3);
    // CPU grid (octree_volume.hpp:42, 147-167) and window width (backproject.hpp:147)
    for (int k = 0; k < 3; ++k) {
        pl->wzero[k] = vpos[k] - vsize[k] / 2.0f;
        pl->wvs[k] = vsize[k] / (float)n[k];
    }
    const float a2 = pl->wvs[0] * pl->wvs[0], b2 = pl->wvs[1] * pl->wvs[1], c2 = pl->wvs[2] * pl->wvs[2];
    const float width = (float)std::sqrt(((double)a2 + (double)b2) + (double)c2);
    pl->half_width = width / 2.0f;
    const std::vector<uint32_t> idx = used_pairs(c, 0);
    std::vector<bpx_pair> pr(idx.size());
    for (size_t i = 0; i < idx.size(); ++i) pr[i] = c.pairs[idx[i]];
    pl->P = (uint32_t)pr.size();
    bool confocal = true;
    for (const auto& p : pr) confocal = confocal && same3(p.laser_point, p.camera_point);
    pl->all_confocal = confocal;
    const size_t nvox = (size_t)n[0] * n[1] * n[2];
    pl->vol.alloc(nvox * (c.cplx ? 8 : 4));
    // exactness precondition: CPU-grid corners
    double corners[8][3];
    for (int m = 0; m < 8; ++m) {
        const double i[3] = {(m & 4) ? (double)(n[0] - 1) : 0.0, (m & 2) ? (double)(n[1] - 1) : 0.0,
                             (m & 1) ? (double)(n[2] - 1) : 0.0};
        for (int k = 0; k < 3; ++k) corners[m][k] = (double)pl->wzero[k] + (double)pl->wvs[k] * i[k];
    }
    const double bound = tap_error_bound(max_total_path(pr.data(), pr.size(), corners) + width, c.t0, c.dT, c.T);
    if (bound >= 0.25)
        fail(BPX_ERR_UNSUPPORTED, "path lengths / t0 too large relative to deltaT for the exact bin lookup");
    pl->inv = (float)(1.0 / (double)c.dT);
    pl->c1_lo = (float)(1.0 - ((double)c.t0 + (double)pl->half_width) / (double)c.dT);
    pl->c1_hi = (float)(1.0 - ((double)c.t0 - (double)pl->half_width) / (double)c.dT);
    if (pl->P == 0 || c.T < 2) return pl.release();
    stage_raw(*pl, c, idx);
    pl->ldh = round_up(c.T, 32);
    pl->H.alloc((size_t)pl->P * pl->ldh * (c.cplx ? 16 : 8));
    pl->Z.alloc(sizeof(float) * c.T);
    pl->Z2.alloc(sizeof(float) * c.T);
    std::vector<float> p8((size_t)pl->P * 8);
    std::memcpy(p8.data(), pr.data(), sizeof(float) * p8.size());
    pl->pair8.alloc(sizeof(float) * p8.size());
    hip_check(hipMemcpy(pl->pair8.ptr, p8.data(), sizeof(float) * p8.size(), hipMemcpyHostToDevice),
              "hipMemcpy(pairs)");
    return pl.release();
}

}  // namespace bpx
