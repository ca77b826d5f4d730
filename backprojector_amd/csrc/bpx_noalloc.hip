// bpx_backprojection_noalloc: the single-tap backprojection of bpx_backprojection on caller-owned
// device buffers, for callers that cannot allocate device memory.  The library's working tables
// (the pair table and the path-length threshold table) live in one fixed file-scope device array,
// `scratch`; its address on each GPU is looked up once (hipGetSymbolAddress, the first call on
// that device) and only that address is passed to kernels and copies.  No hipMalloc, hipFree or
// allocating library call is made on the way to a result.
//
// The gather is the arbitrary-pair kernel reading the caller's unpadded rows directly
// (pairs_kernel, RAW): both legs of the path with the exact square root, the bin decided
// against the capture's path-length thresholds, the same per-update values and Kahan blocks of
// 32 pairs in pair order as the table-driven kernels, so the volume equals bpx_backprojection's.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>

#include "bpx_host.hpp"
#include "bpx_internal.hpp"

namespace {

// 8 MiB per GPU, held until the process ends.  Layout of one call: pair records [P] (32 B
// each) at 0, then the path-length thresholds Z[T + 1] at the next 256-byte boundary.
static __device__ __align__(256) unsigned char scratch[BPX_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once per device).
void* scratch_address() {
    int dev = -1;
    bpx::hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) bpx::fail(BPX_ERR_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        bpx::hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p) bpx::fail(BPX_ERR_DEVICE, "scratch array has no device address");
        g_scratch[dev] = p;
    }
    return g_scratch[dev];
}

bool finite3(const float* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

}  // namespace

using namespace bpx;

extern "C" bpx_status bpx_backprojection_noalloc(const float* transient_chunk, uint64_t transient_size,
                                                 uint32_t T, const bpx_pair* scanned_pairs, uint32_t num_pairs,
                                                 float* voxel_volume, const uint32_t voxels_per_side[3],
                                                 const float volume_zero_pos[3], const float voxel_inc[9],
                                                 float t0, float deltaT, void* hip_stream) {
    clear_error();
    try {
        if (!voxel_volume || !voxels_per_side || !volume_zero_pos || !voxel_inc)
            fail(BPX_ERR_INVALID_ARGUMENT, "NULL volume/grid argument");
        if (num_pairs && (!transient_chunk || !scanned_pairs)) fail(BPX_ERR_INVALID_ARGUMENT, "NULL buffer");
        if (T == 0) fail(BPX_ERR_INVALID_ARGUMENT, "T must be > 0");
        if ((uint64_t)num_pairs * T != transient_size)
            fail(BPX_ERR_INVALID_ARGUMENT, "transient_size != num_pairs * T");
        if (!std::isfinite(t0)) fail(BPX_ERR_INVALID_ARGUMENT, "t0 must be finite");
        if (!(deltaT > 0.0f) || !std::isfinite(deltaT) || !std::isfinite(1.0f / deltaT))
            fail(BPX_ERR_INVALID_ARGUMENT, "deltaT must be a positive normal float");
        const uint32_t* n = voxels_per_side;
        for (int k = 0; k < 3; ++k) {
            if (n[k] == 0) fail(BPX_ERR_INVALID_ARGUMENT, "voxels_per_side must be > 0");
            if (!std::isfinite(volume_zero_pos[k])) fail(BPX_ERR_INVALID_ARGUMENT, "volume_zero_pos must be finite");
        }
        for (int k = 0; k < 9; ++k)
            if (!std::isfinite(voxel_inc[k])) fail(BPX_ERR_INVALID_ARGUMENT, "voxel_inc must be finite");
        const uint64_t nvox = (uint64_t)n[0] * n[1] * n[2];
        if (nvox > (1ull << 32)) fail(BPX_ERR_UNSUPPORTED, "volumes above 2^32 voxels are not supported");
        // scratch capacity: pair records, then the thresholds at a 256-byte boundary
        const uint64_t z_off = ((uint64_t)num_pairs * sizeof(bpx_pair) + 255u) / 256u * 256u;
        const uint64_t need = z_off + 4ull * ((uint64_t)T + 1u);
        if (need > BPX_NOALLOC_SCRATCH_BYTES)
            fail(BPX_ERR_UNSUPPORTED, "capture exceeds the fixed scratch array: num_pairs * 32 + (T + 1) * 4 "
                                      "must stay within BPX_NOALLOC_SCRATCH_BYTES");
        if (4ull * ((uint64_t)T + 1u) > 64u * 1024u) fail(BPX_ERR_UNSUPPORTED, "T must be <= 16383");
        if (4ull * transient_size >= (1ull << 32)) fail(BPX_ERR_UNSUPPORTED, "transient chunk above 4 GiB");
        for (uint32_t p = 0; p < num_pairs; ++p) {
            const bpx_pair& q = scanned_pairs[p];
            if (!finite3(q.camera_point) || !finite3(q.laser_point) || !std::isfinite(q.camera_wall) ||
                !std::isfinite(q.laser_wall))
                fail(BPX_ERR_INVALID_ARGUMENT, "non-finite pair (bpx_backprojection drops such pairs; this entry "
                                               "point takes finite pairs only)");
        }
        if (single_tap_error_bound(scanned_pairs, num_pairs, n, volume_zero_pos, voxel_inc, t0, deltaT, T) >= 0.25)
            fail(BPX_ERR_UNSUPPORTED, "path lengths / t0 too large relative to deltaT for the exact bin lookup");
        if (!device_ready()) fail(BPX_ERR_NO_DEVICE, "no gfx950 device available");
        char* base = static_cast<char*>(scratch_address());
        hipStream_t s = static_cast<hipStream_t>(hip_stream);
        if (num_pairs == 0) {
            hip_check(launch_fill(voxel_volume, nvox, 0.0f, s), "fill kernel");
        } else {
            float* Z = reinterpret_cast<float*>(base + z_off);
            hip_check(hipMemcpyAsync(base, scanned_pairs, sizeof(bpx_pair) * num_pairs, hipMemcpyHostToDevice, s),
                      "hipMemcpyAsync(pairs -> scratch)");
            hip_check(launch_ztable(0, t0, deltaT, 0.0f, T + 1u, Z, s), "threshold table kernel");
            PairsArgs a{};
            a.H = transient_chunk;
            a.pairs = reinterpret_cast<const float*>(base);
            a.Z = Z;
            a.out = voxel_volume;
            a.P = num_pairs;
            a.Tc = T;
            a.ldh = T;
            a.h_bytes = (uint32_t)(4ull * transient_size);
            std::memcpy(a.g.n, n, sizeof(uint32_t) * 3);
            std::memcpy(a.g.zero, volume_zero_pos, sizeof(float) * 3);
            std::memcpy(a.g.inc, voxel_inc, sizeof(float) * 9);
            a.t = make_tiles(n, 4);
            a.inv = (float)(1.0 / (double)deltaT);
            a.c1 = (float)(1.0 - (double)t0 / (double)deltaT);
            a.scale = 1.0f;
            a.kahan_block = 32;
            hip_check(launch_pairs_raw(a, s), "pairs gather kernel (raw rows)");
        }
        // the scratch array and the caller's host pair records are free again once this returns
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
        return BPX_OK;
    } catch (const Failure& e) {
        return set_error(e.status, e.msg);
    }
}
