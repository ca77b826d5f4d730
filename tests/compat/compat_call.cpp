// Calls the reference-signature C++ ABI exactly as the reference's host pipeline does
// (backproject.hpp:883-891): a confocal 8x8 capture with integer sample codes into a 6x5x4
// volume.  Prints the volume as text (one value per line) for the test to compare.
#include <cmath>
#include <cstdio>
#include <vector>

#include "backproject_cuda_compat.hpp"

int main() {
    const uint32_t side = 8, T = 1500;
    std::vector<CameraLaserPair> pairs;
    for (uint32_t x = 0; x < side; ++x)
        for (uint32_t y = 0; y < side; ++y) {
            CameraLaserPair p{};
            const float wx = -1.0f + 2.0f * x / (side - 1), wy = -1.0f + 2.0f * y / (side - 1);
            p.camera_point[0] = p.laser_point[0] = wx;
            p.camera_point[1] = p.laser_point[1] = wy;
            const double d = std::sqrt((double)wx * wx + (double)(wy + 1.0f) * (wy + 1.0f) + 0.25);
            p.camera_wall = p.laser_wall = (float)d;
            pairs.push_back(p);
        }
    std::vector<float> H(pairs.size() * T);
    for (size_t p = 0; p < pairs.size(); ++p)
        for (uint32_t b = 0; b < T; ++b) H[p * T + b] = (float)((7 * b + 13 * p) % 61 + 1);
    const uint32_t vps[3] = {6, 5, 4};
    const float zero[3] = {-0.5f, -0.5f, 0.3f};
    const float inc[9] = {0.2f, 0, 0, 0, 0.25f, 0, 0, 0, 1.0f / 3.0f};
    std::vector<float> vol(6 * 5 * 4, -1.0f);
    call_cuda_backprojection(H.data(), (uint32_t)H.size(), T, pairs, vol.data(), vps, zero, inc,
                             2.0f, 0.0025f);
    for (float v : vol) std::printf("%.1f\n", v);
    return 0;
}
