// C++ compatibility declarations: the reference's device ABI (backproject_cuda.hpp:9-49) with
// identical names and signatures, implemented by libbpx on gfx950.  A caller written against
// the reference header (its backproject.hpp host pipeline, its main.cpp) compiles against this
// header and links libbpx.so unchanged.
//
// Differences from the reference: errors throw std::runtime_error (the reference prints and
// calls exit(1)); nothing is printed; every pair is used (the reference launcher silently
// drops the last num_pairs mod min(256, num_pairs) pairs, backproject_cuda.cu:150, 418).
// call_cuda_octree_backprojection runs this library's coarse-to-fine octree (the reference's
// own GPU octree prints "GPU OCTREE BP NOT WORKING YET", backproject_cuda.cu:513).
#ifndef BACKPROJECT_CUDA_COMPAT_HPP
#define BACKPROJECT_CUDA_COMPAT_HPP

#include <cstdint>
#include <complex>
#include <vector>

#ifndef BACKPROJECT_CUDA
#define BACKPROJECT_CUDA
struct CameraLaserPair {
    float camera_point[3], laser_point[3];
    float camera_wall, laser_wall;
};
#endif

void call_cuda_backprojection(const float* transient_chunk, uint32_t transient_size, uint32_t T,
                              const std::vector<CameraLaserPair> scanned_pairs,
                              float* voxel_volume, const uint32_t* voxels_per_side,
                              const float* volume_zero_pos, const float* voxel_inc, float t0,
                              float deltaT);

void call_cuda_octree_backprojection(const float* transient_chunk, uint32_t transient_size,
                                     uint32_t T, const std::vector<CameraLaserPair> scanned_pairs,
                                     float* voxel_volume, const uint32_t* voxels_per_side,
                                     const float* volume_zero_pos, const float* voxel_inc,
                                     float t0, float deltaT);

void call_cuda_complex_backprojection(const std::complex<float>* transient_chunk,
                                      uint32_t transient_size, uint32_t T,
                                      const std::vector<CameraLaserPair> scanned_pairs,
                                      std::complex<float>* voxel_volume,
                                      const uint32_t* voxels_per_side,
                                      const float* volume_zero_pos, const float* voxel_inc,
                                      float t0, float deltaT);

#endif
