// extern "C" entry points of libbpx (declared in include/bpx.h).  Every entry point catches
// all failures and converts them to a bpx_status plus a thread-local message.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>

#include <vector>

#include "bpx_host.hpp"

namespace bpx {

static thread_local std::string g_error;

bpx_status set_error(bpx_status s, const std::string& msg) {
    g_error = msg;
    return s;
}
void clear_error() { g_error.clear(); }
void fail(bpx_status s, const std::string& msg) { throw Failure{s, msg}; }
void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess)
        fail(BPX_ERR_DEVICE, std::string(what) + ": " + hipGetErrorName(e) + " (" +
                                 hipGetErrorString(e) + ")");
}

static int g_device = -2;  // -2 unknown, -1 none
static std::mutex g_dev_mu;

static int find_device() {
    std::lock_guard<std::mutex> lk(g_dev_mu);
    if (g_device != -2) return g_device;
    g_device = -1;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return g_device;
    }
    for (int d = 0; d < n; ++d) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, d) != hipSuccess) continue;
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) == 0) {
            g_device = d;
            break;
        }
    }
    return g_device;
}

bool device_ready() {
    const int d = find_device();
    if (d < 0) return false;
    return hipSetDevice(d) == hipSuccess;
}

template <class F>
static bpx_status guarded(F&& f) {
    clear_error();
    try {
        f();
        return BPX_OK;
    } catch (const Failure& e) {
        return set_error(e.status, e.msg);
    } catch (const std::bad_alloc&) {
        return set_error(BPX_ERR_OUT_OF_MEMORY, "host allocation failed");
    } catch (const std::exception& e) {
        return set_error(BPX_ERR_INVALID_ARGUMENT, e.what());
    }
}

void run_plan_to_host(PlanImpl* raw, float* out) {
    std::unique_ptr<PlanImpl> pl(raw);
    pl->execute(pl->stream);
    const size_t nvox = (size_t)pl->g.n[0] * pl->g.n[1] * pl->g.n[2];
    hip_check(hipMemcpyAsync(out, pl->vol.ptr, nvox * (pl->cplx ? 8 : 4), hipMemcpyDeviceToHost,
                             pl->stream),
              "hipMemcpyAsync(volume)");
    hip_check(hipStreamSynchronize(pl->stream), "hipStreamSynchronize");
}

}  // namespace bpx

using namespace bpx;

extern "C" {

const char* bpx_last_error(void) { return g_error.c_str(); }
const char* bpx_version(void) { return "bpx 0.1.0 (gfx950)"; }
int bpx_abi_version(void) { return BPX_ABI_VERSION; }
int bpx_device_count(void) { return find_device() >= 0 ? 1 : 0; }

bpx_status bpx_backprojection_ex(const float* transient_chunk, uint64_t transient_size, uint32_t T,
                                 const bpx_pair* scanned_pairs, uint32_t num_pairs,
                                 float* voxel_volume, const uint32_t voxels_per_side[3],
                                 const float volume_zero_pos[3], const float voxel_inc[9],
                                 float t0, float deltaT, int complex_samples,
                                 const bpx_options* options) {
    return guarded([&] {
        if (!voxel_volume || !voxels_per_side || !volume_zero_pos || !voxel_inc)
            fail(BPX_ERR_INVALID_ARGUMENT, "NULL volume/grid argument");
        CaptureIn c{transient_chunk, transient_size, T, scanned_pairs, num_pairs, t0, deltaT,
                    complex_samples != 0};
        run_plan_to_host(make_plan_single_tap(c, voxels_per_side, volume_zero_pos, voxel_inc, options),
                         voxel_volume);
    });
}

bpx_status bpx_backprojection(const float* transient_chunk, uint64_t transient_size, uint32_t T,
                              const bpx_pair* scanned_pairs, uint32_t num_pairs,
                              float* voxel_volume, const uint32_t voxels_per_side[3],
                              const float volume_zero_pos[3], const float voxel_inc[9], float t0,
                              float deltaT) {
    return bpx_backprojection_ex(transient_chunk, transient_size, T, scanned_pairs, num_pairs,
                                 voxel_volume, voxels_per_side, volume_zero_pos, voxel_inc, t0,
                                 deltaT, 0, nullptr);
}

bpx_status bpx_complex_backprojection(const float* transient_chunk, uint64_t transient_size,
                                      uint32_t T, const bpx_pair* scanned_pairs,
                                      uint32_t num_pairs, float* voxel_volume,
                                      const uint32_t voxels_per_side[3],
                                      const float volume_zero_pos[3], const float voxel_inc[9],
                                      float t0, float deltaT) {
    return bpx_backprojection_ex(transient_chunk, transient_size, T, scanned_pairs, num_pairs,
                                 voxel_volume, voxels_per_side, volume_zero_pos, voxel_inc, t0,
                                 deltaT, 1, nullptr);
}

bpx_status bpx_window_backprojection(const float* transient_chunk, uint64_t transient_size,
                                     uint32_t T, const bpx_pair* scanned_pairs,
                                     uint32_t num_pairs, float* voxel_volume,
                                     const uint32_t voxels_per_side[3],
                                     const float volume_position[3], const float volume_size[3],
                                     float t0, float deltaT, int complex_samples) {
    return guarded([&] {
        if (!voxel_volume || !voxels_per_side || !volume_position || !volume_size)
            fail(BPX_ERR_INVALID_ARGUMENT, "NULL volume/grid argument");
        CaptureIn c{transient_chunk, transient_size, T, scanned_pairs, num_pairs, t0, deltaT,
                    complex_samples != 0};
        run_plan_to_host(make_plan_window(c, voxels_per_side, volume_position, volume_size),
                         voxel_volume);
    });
}

bpx_status bpx_phasor_pulse(const float* transient_in, uint64_t rows, uint32_t T, float wavelength,
                            float deltaT, uint32_t times, float sigma, float* complex_out) {
    return guarded([&] {
        if ((!transient_in || !complex_out) && rows) fail(BPX_ERR_INVALID_ARGUMENT, "NULL buffer");
        if (!(deltaT > 0.0f) || !std::isfinite(deltaT)) fail(BPX_ERR_INVALID_ARGUMENT, "deltaT must be > 0");
        if (!(wavelength > 0.0f) || !std::isfinite(wavelength))
            fail(BPX_ERR_INVALID_ARGUMENT, "wavelength must be > 0");
        std::vector<float> c, s;
        if (pipeline::phasor_taps(wavelength, deltaT, times, sigma, c, s) <= 0)
            fail(BPX_ERR_INVALID_ARGUMENT, "the phasor pulse has no taps for these parameters");
        if (!device_ready()) fail(BPX_ERR_NO_DEVICE, "no gfx950 device available");
        phasor_filter_host(transient_in, rows, T, c, s, complex_out);
    });
}

bpx_status bpx_plan_create(const float* transient_chunk, uint64_t transient_size, uint32_t T,
                           const bpx_pair* scanned_pairs, uint32_t num_pairs,
                           const uint32_t voxels_per_side[3], const float volume_zero_pos[3],
                           const float voxel_inc[9], float t0, float deltaT, int complex_samples,
                           const bpx_options* options, bpx_plan* plan_out) {
    return guarded([&] {
        if (!plan_out || !voxels_per_side || !volume_zero_pos || !voxel_inc)
            fail(BPX_ERR_INVALID_ARGUMENT, "NULL argument");
        *plan_out = nullptr;
        CaptureIn c{transient_chunk, transient_size, T, scanned_pairs, num_pairs, t0, deltaT,
                    complex_samples != 0};
        auto p = std::make_unique<bpx_plan_s>();
        p->impl.reset(make_plan_single_tap(c, voxels_per_side, volume_zero_pos, voxel_inc, options));
        *plan_out = p.release();
    });
}

static void fill_octree_stats(const PlanImpl& pl, bpx_octree_stats* s) {
    std::memset(s, 0, sizeof(*s));
    if (!pl.oct) return;
    const OctreeState& st = *pl.oct;
    s->levels = st.L;
    for (uint32_t l = 0; l <= st.L && l < st.level_active.size(); ++l) s->active[l] = st.level_active[l];
    for (uint32_t l = 0; l < st.L && 5 * l + 4 < st.level_stats.size(); ++l) {
        s->running_max[l] = st.level_stats[5 * l + 0];
        s->bound_max[l] = st.level_stats[5 * l + 1];
        s->bound_norm[l] = std::sqrt(st.level_stats[5 * l + 2]);
        s->threshold[l] = st.level_stats[5 * l + 4];
    }
    uint64_t b = 0;
    for (uint32_t l = 0; l < st.L && l < st.level_active.size(); ++l) b += (uint64_t)st.level_active[l] * st.coarse_P;
    s->bound_updates = b;
    s->fine_updates = st.level_active.size() > st.L ? (uint64_t)st.level_active[st.L] * pl.P : 0;
    for (uint32_t v : st.fallback_voxels) s->fallback_voxels += v;
    s->coarse_pairs = st.coarse_P;
}

bpx_status bpx_octree_backprojection(const float* transient_chunk, uint64_t transient_size,
                                     uint32_t T, const bpx_pair* scanned_pairs, uint32_t num_pairs,
                                     float* voxel_volume, const uint32_t voxels_per_side[3],
                                     const float volume_zero_pos[3], const float voxel_inc[9],
                                     float t0, float deltaT, const bpx_octree_options* options,
                                     bpx_octree_stats* stats) {
    return guarded([&] {
        if (!voxel_volume || !voxels_per_side || !volume_zero_pos || !voxel_inc)
            fail(BPX_ERR_INVALID_ARGUMENT, "NULL volume/grid argument");
        CaptureIn c{transient_chunk, transient_size, T, scanned_pairs, num_pairs, t0, deltaT, false};
        std::unique_ptr<PlanImpl> pl(make_plan_octree(c, voxels_per_side, volume_zero_pos, voxel_inc, options));
        pl->execute(pl->stream);
        const size_t nvox = (size_t)pl->g.n[0] * pl->g.n[1] * pl->g.n[2];
        hip_check(hipMemcpyAsync(voxel_volume, pl->vol.ptr, nvox * 4, hipMemcpyDeviceToHost, pl->stream),
                  "hipMemcpyAsync(volume)");
        hip_check(hipStreamSynchronize(pl->stream), "hipStreamSynchronize");
        if (stats) fill_octree_stats(*pl, stats);
    });
}

bpx_status bpx_octree_plan_create(const float* transient_chunk, uint64_t transient_size, uint32_t T,
                                  const bpx_pair* scanned_pairs, uint32_t num_pairs,
                                  const uint32_t voxels_per_side[3], const float volume_zero_pos[3],
                                  const float voxel_inc[9], float t0, float deltaT,
                                  const bpx_octree_options* options, bpx_plan* plan_out) {
    return guarded([&] {
        if (!plan_out || !voxels_per_side || !volume_zero_pos || !voxel_inc)
            fail(BPX_ERR_INVALID_ARGUMENT, "NULL argument");
        *plan_out = nullptr;
        CaptureIn c{transient_chunk, transient_size, T, scanned_pairs, num_pairs, t0, deltaT, false};
        auto p = std::make_unique<bpx_plan_s>();
        p->impl.reset(make_plan_octree(c, voxels_per_side, volume_zero_pos, voxel_inc, options));
        *plan_out = p.release();
    });
}

bpx_status bpx_octree_plan_stats(bpx_plan plan, bpx_octree_stats* stats) {
    return guarded([&] {
        if (!plan || !stats) fail(BPX_ERR_INVALID_ARGUMENT, "NULL argument");
        if (!plan->impl->oct) fail(BPX_ERR_INVALID_ARGUMENT, "not an octree plan");
        plan_wait(*plan->impl);
        fill_octree_stats(*plan->impl, stats);
    });
}

bpx_status bpx_octree_plan_set_options(bpx_plan plan, const bpx_octree_options* options) {
    return guarded([&] {
        if (!plan || !options) fail(BPX_ERR_INVALID_ARGUMENT, "NULL argument");
        if (!plan->impl->oct) fail(BPX_ERR_INVALID_ARGUMENT, "not an octree plan");
        plan_wait(*plan->impl);
        octree_set_options(*plan->impl->oct, *options);
    });
}

bpx_status bpx_plan_execute(bpx_plan plan, void* hip_stream) {
    return guarded([&] {
        if (!plan) fail(BPX_ERR_INVALID_ARGUMENT, "NULL plan");
        PlanImpl& pl = *plan->impl;
        const hipStream_t s = hip_stream ? (hipStream_t)hip_stream : pl.stream;
        // Executions of one plan share its buffers: order this one after the work queued on the
        // plan's own stream and after the previous execution, whichever stream that used.
        if (s != pl.stream) {
            hip_check(hipEventRecord(pl.ready, pl.stream), "hipEventRecord");
            hip_check(hipStreamWaitEvent(s, pl.ready, 0), "hipStreamWaitEvent");
        }
        hip_check(hipStreamWaitEvent(s, pl.done, 0), "hipStreamWaitEvent");
        pl.execute(s);
        hip_check(hipEventRecord(pl.done, s), "hipEventRecord");
    });
}

bpx_status bpx_plan_synchronize(bpx_plan plan) {
    return guarded([&] {
        if (!plan) fail(BPX_ERR_INVALID_ARGUMENT, "NULL plan");
        plan_wait(*plan->impl);
    });
}

bpx_status bpx_plan_read_volume(bpx_plan plan, float* voxel_volume) {
    return guarded([&] {
        if (!plan || !voxel_volume) fail(BPX_ERR_INVALID_ARGUMENT, "NULL argument");
        PlanImpl& pl = *plan->impl;
        const size_t nvox = (size_t)pl.g.n[0] * pl.g.n[1] * pl.g.n[2];
        plan_wait(pl);
        hip_check(hipMemcpy(voxel_volume, pl.vol.ptr, nvox * (pl.cplx ? 8 : 4), hipMemcpyDeviceToHost),
                  "hipMemcpy(volume)");
    });
}

bpx_status bpx_row_windows(const bpx_pair* pairs, uint32_t num_pairs, const uint32_t voxels_per_side[3],
                           const float volume_zero_pos[3], const float voxel_inc[9], float t0, float deltaT,
                           uint32_t Tc, uint32_t* out) {
    return guarded([&] {
        if ((num_pairs && (!pairs || !out)) || !voxels_per_side || !volume_zero_pos || !voxel_inc)
            fail(BPX_ERR_INVALID_ARGUMENT, "NULL argument");
        for (int k = 0; k < 3; ++k)
            if (voxels_per_side[k] == 0) fail(BPX_ERR_INVALID_ARGUMENT, "voxels_per_side must be positive");
        if (!(deltaT > 0.0f)) fail(BPX_ERR_INVALID_ARGUMENT, "deltaT must be positive");
        const std::vector<uint32_t> w =
            confocal_row_windows(pairs, num_pairs, voxels_per_side, volume_zero_pos, voxel_inc, t0, deltaT, Tc);
        std::copy(w.begin(), w.end(), out);
    });
}

bpx_status bpx_plan_read_thresholds(bpx_plan plan, float* out, uint64_t capacity) {
    return guarded([&] {
        if (!plan || !out) fail(BPX_ERR_INVALID_ARGUMENT, "NULL argument");
        const PlanImpl& pl = *plan->impl;
        if (pl.path != kPathConfocal && pl.path != kPathOctree)
            fail(BPX_ERR_UNSUPPORTED, "threshold tables belong to confocal plans");
        const uint64_t n = (uint64_t)pl.P * (pl.T + 1u);
        if (capacity < n) fail(BPX_ERR_INVALID_ARGUMENT, "output holds fewer than pairs_used * (Tc + 1) floats");
        if (n == 0) return;
        plan_wait(pl);
        const DeviceBuffer& src = pl.fused ? pl.H : pl.S;
        const int fmt = staged_format(pl);
        const size_t stride = pl.fused ? (size_t)fused_elem_bytes(fmt) / 4 : 1;
        const size_t pitch = pl.fused ? (size_t)pl.ldh * stride : pl.lds;
        const size_t off = pl.fused ? (size_t)fused_s_offset(fmt) : 0;
        std::vector<float> host(src.bytes / 4);
        hip_check(hipMemcpy(host.data(), src.ptr, host.size() * 4, hipMemcpyDeviceToHost), "hipMemcpy(thresholds)");
        // rows staged over windows: the entries outside them come from the same search, run
        // here over the whole table
        std::vector<float> full;
        if (pl.fused && !pl.win_host.empty()) {
            DeviceBuffer tmp;
            tmp.alloc(sizeof(float) * n);
            hip_check(launch_stable(static_cast<const float*>(pl.pair8.ptr), pl.P, pl.T,
                                    static_cast<const float*>(pl.Z.ptr), static_cast<float*>(tmp.ptr), pl.T + 1u, 1u,
                                    0u, pl.stream),
                      "threshold table kernel");
            hip_check(hipStreamSynchronize(pl.stream), "hipStreamSynchronize");
            full.resize(n);
            hip_check(hipMemcpy(full.data(), tmp.ptr, sizeof(float) * n, hipMemcpyDeviceToHost), "hipMemcpy(thresholds)");
        }
        for (uint64_t p = 0; p < pl.P; ++p)
            for (uint64_t j = 0; j <= pl.T; ++j) {
                const bool staged = full.empty() || (j >= pl.win_host[2 * p] && j < pl.win_host[2 * p + 1]);
                out[p * (pl.T + 1u) + j] = staged ? host[p * pitch + j * stride + off] : full[p * (pl.T + 1u) + j];
            }
    });
}

bpx_status bpx_plan_info(bpx_plan plan, int32_t* path, int32_t* lane_voxels, uint32_t* pairs_used,
                         uint32_t* Tc, float* scale) {
    return guarded([&] {
        if (!plan) fail(BPX_ERR_INVALID_ARGUMENT, "NULL plan");
        const PlanImpl& pl = *plan->impl;
        if (path) *path = pl.path;
        if (lane_voxels) *lane_voxels = pl.lds_tiles ? lds_shape(pl.variant, pl.format).kz : pl.K;
        if (pairs_used) *pairs_used = pl.P;
        if (Tc) *Tc = pl.T;
        if (scale) *scale = pl.scale;
    });
}

bpx_status bpx_plan_row_storage(bpx_plan plan, uint32_t* storage) {
    return guarded([&] {
        if (!plan || !storage) fail(BPX_ERR_INVALID_ARGUMENT, "NULL argument");
        const int f = staged_format(*plan->impl);
        *storage = f == kF16 ? BPX_STORAGE_F16 : f == kBF16 ? BPX_STORAGE_BF16 : BPX_STORAGE_F32;
    });
}

bpx_status bpx_plan_kernel_name(bpx_plan plan, char* out, uint64_t capacity) {
    return guarded([&] {
        if (!plan || !out || capacity == 0) fail(BPX_ERR_INVALID_ARGUMENT, "NULL argument");
        const PlanImpl& pl = *plan->impl;
        std::string name;
        if (pl.lds_tiles) {
            const LdsShape sh = lds_shape(pl.variant, pl.format);
            name = "confocal_lds_kernel<KZ=" + std::to_string(sh.kz) + ", waves=" + std::to_string(sh.nw) +
                   "> (16x16x" + std::to_string(sh.depth()) + " tiles)";
            if (pl.format == kC64) name += " x2 (real and imaginary passes)";
        } else {
            static const char* paths[] = {"confocal_kernel", "grid_kernel", "pairs_kernel", "window_kernel",
                                          "octree kernels"};
            name = std::string(paths[pl.path]) + "<K=" + std::to_string(pl.K) + ">";
        }
        std::snprintf(out, (size_t)capacity, "%s", name.c_str());
    });
}

bpx_status bpx_plan_benchmark(bpx_plan plan, int32_t warmup, int32_t iterations, bpx_timing* out) {
    return guarded([&] {
        if (!plan || !out || iterations <= 0 || warmup < 0) fail(BPX_ERR_INVALID_ARGUMENT, "bad arguments");
        PlanImpl& pl = *plan->impl;
        plan_wait(pl);
        for (int i = 0; i < warmup; ++i) pl.execute(pl.stream);
        hip_check(hipStreamSynchronize(pl.stream), "hipStreamSynchronize");
        double table_ms = 0, gather_ms = 0;
        const auto t_start = std::chrono::steady_clock::now();
        for (int i = 0; i < iterations; ++i) {
            pl.execute(pl.stream);
            if (pl.P == 0) continue;
            hip_check(hipEventSynchronize(pl.ev[3]), "hipEventSynchronize");
            float a = 0, b = 0;
            hip_check(hipEventElapsedTime(&a, pl.ev[1], pl.ev[2]), "hipEventElapsedTime");
            hip_check(hipEventElapsedTime(&b, pl.ev[2], pl.ev[3]), "hipEventElapsedTime");
            table_ms += a;
            gather_ms += b;
        }
        hip_check(hipStreamSynchronize(pl.stream), "hipStreamSynchronize");
        const auto t_end = std::chrono::steady_clock::now();
        out->iterations = iterations;
        out->total_ms = std::chrono::duration<double, std::milli>(t_end - t_start).count();
        out->per_exec_ms = out->total_ms / iterations;
        out->table_kernel_ms = table_ms / iterations;
        out->gather_kernel_ms = gather_ms / iterations;
    });
}

bpx_status bpx_plan_tune(bpx_plan plan, int32_t reps, int32_t* variant) {
    return guarded([&] {
        if (!plan) fail(BPX_ERR_INVALID_ARGUMENT, "NULL plan");
        PlanImpl& pl = *plan->impl;
        plan_wait(pl);
        if (pl.path == kPathConfocal && pl.fused && pl.P) {
            reps = std::max(reps, 1);
            hipEvent_t t0 = nullptr, t1 = nullptr;
            hip_check(hipEventCreate(&t0), "hipEventCreate");
            if (hipEventCreate(&t1) != hipSuccess) {
                (void)hipEventDestroy(t0);
                fail(BPX_ERR_DEVICE, "hipEventCreate failed");
            }
            const int initial = pl.variant;
            int best = initial;
            float best_ms = INFINITY;
            try {
                for (int v = 3; v < variant_count(); ++v) {
                    set_fused_variant(pl, v);
                    if (v >= kLdsVariant && !pl.lds_tiles) continue;   // would run as a texture-path shape
                    pl.execute(pl.stream);                             // warm-up
                    hip_check(hipEventRecord(t0, pl.stream), "hipEventRecord");
                    for (int r = 0; r < reps; ++r) pl.execute(pl.stream);
                    hip_check(hipEventRecord(t1, pl.stream), "hipEventRecord");
                    hip_check(hipEventSynchronize(t1), "hipEventSynchronize");
                    float ms = 0;
                    hip_check(hipEventElapsedTime(&ms, t0, t1), "hipEventElapsedTime");
                    if (ms < best_ms) {
                        best_ms = ms;
                        best = v;
                    }
                }
                set_fused_variant(pl, best);
            } catch (...) {
                (void)hipEventDestroy(t0);
                (void)hipEventDestroy(t1);
                set_fused_variant(pl, initial);
                throw;
            }
            (void)hipEventDestroy(t0);
            (void)hipEventDestroy(t1);
        }
        if (variant) *variant = pl.variant;
    });
}

void bpx_plan_destroy(bpx_plan plan) {
    if (!plan) return;
    (void)hipStreamSynchronize(plan->impl->stream);
    if (plan->impl->done) (void)hipEventSynchronize(plan->impl->done);
    delete plan;
}

bpx_status bpx_copy_benchmark(uint64_t bytes, int32_t iterations, double* mean_ms) {
    return guarded([&] {
        if (!mean_ms || iterations <= 0 || bytes < 16) fail(BPX_ERR_INVALID_ARGUMENT, "bad arguments");
        if (!device_ready()) fail(BPX_ERR_NO_DEVICE, "no gfx950 device available");
        bytes &= ~(uint64_t)15;
        DeviceBuffer a, b;
        a.alloc(bytes);
        b.alloc(bytes);
        hip_check(hipMemset(a.ptr, 0, bytes), "hipMemset");
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "hipEventCreate");
        hip_check(hipEventCreate(&e1), "hipEventCreate");
        hip_check(launch_copy(b.ptr, a.ptr, bytes, nullptr), "copy kernel");  // warm-up
        double total = 0;
        for (int i = 0; i < iterations; ++i) {
            hip_check(hipEventRecord(e0, nullptr), "hipEventRecord");
            hip_check(launch_copy(b.ptr, a.ptr, bytes, nullptr), "copy kernel");
            hip_check(hipEventRecord(e1, nullptr), "hipEventRecord");
            hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
            total += ms;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *mean_ms = total / iterations;
    });
}

bpx_status bpx_render_transient(const bpx_pair* pairs, uint32_t num_pairs, uint32_t T, float t0,
                                float deltaT, const float* samples, uint64_t num_samples,
                                float peak_count, int poisson, uint64_t seed, float* out) {
    return guarded([&] {
        if ((!pairs && num_pairs) || (!samples && num_samples) || !out)
            fail(BPX_ERR_INVALID_ARGUMENT, "NULL buffer");
        if (!(deltaT > 0.0f)) fail(BPX_ERR_INVALID_ARGUMENT, "deltaT must be > 0");
        render_transient(pairs, num_pairs, T, t0, deltaT, samples, num_samples, peak_count,
                         poisson != 0, seed, out);
    });
}

}  // extern "C"
