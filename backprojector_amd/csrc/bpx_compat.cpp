// The reference's C++ device ABI (backproject_cuda.hpp:19-49) implemented over the C ABI.
#include <stdexcept>
#include <string>

#include "../../include/backproject_cuda_compat.hpp"
#include "../../include/bpx.h"

static_assert(sizeof(CameraLaserPair) == sizeof(bpx_pair), "CameraLaserPair layout");

static void throw_if(bpx_status s, const char* what) {
    if (s != BPX_OK) throw std::runtime_error(std::string(what) + ": " + bpx_last_error());
}

static uint64_t pairs_times_T(const std::vector<CameraLaserPair>& p, uint32_t T) {
    return (uint64_t)p.size() * T;
}

void call_cuda_backprojection(const float* transient_chunk, uint32_t transient_size, uint32_t T,
                              const std::vector<CameraLaserPair> scanned_pairs,
                              float* voxel_volume, const uint32_t* voxels_per_side,
                              const float* volume_zero_pos, const float* voxel_inc, float t0,
                              float deltaT) {
    if ((uint64_t)transient_size != pairs_times_T(scanned_pairs, T))
        throw std::runtime_error("call_cuda_backprojection: transient_size != pairs * T");
    throw_if(bpx_backprojection(transient_chunk, transient_size, T,
                                reinterpret_cast<const bpx_pair*>(scanned_pairs.data()),
                                (uint32_t)scanned_pairs.size(), voxel_volume, voxels_per_side,
                                volume_zero_pos, voxel_inc, t0, deltaT),
             "call_cuda_backprojection");
}

// The reference's GPU octree is non-functional (backproject_cuda.cu:513); this runs the
// library's coarse-to-fine octree with up to 4 refinements (as many as every side's power of
// two allows), falling back to the dense reconstruction when a side is odd.
void call_cuda_octree_backprojection(const float* transient_chunk, uint32_t transient_size,
                                     uint32_t T, const std::vector<CameraLaserPair> scanned_pairs,
                                     float* voxel_volume, const uint32_t* voxels_per_side,
                                     const float* volume_zero_pos, const float* voxel_inc,
                                     float t0, float deltaT) {
    if ((uint64_t)transient_size != pairs_times_T(scanned_pairs, T))
        throw std::runtime_error("call_cuda_octree_backprojection: transient_size != pairs * T");
    uint32_t levels = 0;
    while (levels < 4 && voxels_per_side[0] % (2u << levels) == 0 &&
           voxels_per_side[1] % (2u << levels) == 0 && voxels_per_side[2] % (2u << levels) == 0)
        ++levels;
    bool confocal = true;
    for (const auto& p : scanned_pairs)
        for (int k = 0; k < 3; ++k) confocal = confocal && p.laser_point[k] == p.camera_point[k];
    if (levels == 0 || !confocal) {
        call_cuda_backprojection(transient_chunk, transient_size, T, scanned_pairs, voxel_volume,
                                 voxels_per_side, volume_zero_pos, voxel_inc, t0, deltaT);
        return;
    }
    const bpx_octree_options o{levels, 0.25f, 0.5f, 0u, {}};
    throw_if(bpx_octree_backprojection(transient_chunk, transient_size, T,
                                       reinterpret_cast<const bpx_pair*>(scanned_pairs.data()),
                                       (uint32_t)scanned_pairs.size(), voxel_volume, voxels_per_side,
                                       volume_zero_pos, voxel_inc, t0, deltaT, &o, nullptr),
             "call_cuda_octree_backprojection");
}

void call_cuda_complex_backprojection(const std::complex<float>* transient_chunk,
                                      uint32_t transient_size, uint32_t T,
                                      const std::vector<CameraLaserPair> scanned_pairs,
                                      std::complex<float>* voxel_volume,
                                      const uint32_t* voxels_per_side,
                                      const float* volume_zero_pos, const float* voxel_inc,
                                      float t0, float deltaT) {
    if ((uint64_t)transient_size != pairs_times_T(scanned_pairs, T))
        throw std::runtime_error("call_cuda_complex_backprojection: transient_size != pairs * T");
    throw_if(bpx_complex_backprojection(reinterpret_cast<const float*>(transient_chunk), transient_size,
                                        T, reinterpret_cast<const bpx_pair*>(scanned_pairs.data()),
                                        (uint32_t)scanned_pairs.size(),
                                        reinterpret_cast<float*>(voxel_volume), voxels_per_side,
                                        volume_zero_pos, voxel_inc, t0, deltaT),
             "call_cuda_complex_backprojection");
}
