/*
 * bp_oracle.c -- CPU restatement of the reference NLOS backprojection, used ONLY as a
 * checker (tests/, __graft_entry__.smoke(), bench.py's cpu_baseline leg).  It is test
 * infrastructure: the product library (backprojector_amd/libbpx.so) never links, loads
 * or calls it.
 *
 * Every function follows a reference file:line (paths relative to the reference repo
 * mjgalindo/backprojector).  Arithmetic is IEEE binary32/binary64 in exactly the order
 * the reference writes it; this file MUST be compiled with -ffp-contract=off and without
 * -ffast-math (oracle/Makefile does so) so that no multiply-add is fused.
 *
 * Two reference formulas exist and both are restated:
 *   A.1 (GPU path)  backproject_cuda.cu:25-46 (distance / compute_distance),
 *                   :144-148 (voxel position), :150-162 (single-tap window, footprint 0),
 *                   :165-173 (summation order).  The reference calls CUDA's norm3df; that
 *                   function's internals are not public and are not reproduced: distances are
 *                   sqrtf((a*a + b*b) + c*c) with an IEEE correctly-rounded sqrtf.
 *   A.2 (CPU path)  backproject.hpp:42-53 (double distance), :75-96 (accumulate_radiance),
 *                   :138-216 (classic_backprojection), octree_volume.hpp:42,147-167 (grid).
 *
 * Where a result is an integer bin the rounding is C99 roundf (half away from zero), as
 * CUDA's round() and C++ round() are.  Out-of-range float->int conversions (undefined in C)
 * are resolved by comparing in float before converting: a rounded bin r contributes iff
 * 0 <= r <= T-1; NaN or infinite path lengths contribute nothing.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <limits.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Same layout as the reference's struct CameraLaserPair (backproject_cuda.hpp:9-13). */
typedef struct {
    float camera_point[3];
    float laser_point[3];
    float camera_wall;
    float laser_wall;
} or_pair;

/* ------------------------------------------------------------------------------------ */
/* A.1 -- the GPU path formula                                                           */
/* ------------------------------------------------------------------------------------ */

/* backproject_cuda.cu:144-148: v[k] = zero[k] + inc[k]*x + inc[3+k]*y + inc[6+k]*z,
 * evaluated left to right; the uint32 voxel index is converted to float. */
void or_a1_voxel_position(const float zero[3], const float inc[9],
                          uint32_t ix, uint32_t iy, uint32_t iz, float v[3])
{
    const float fx = (float)ix, fy = (float)iy, fz = (float)iz;
    for (int k = 0; k < 3; ++k) {
        float a = zero[k] + inc[k] * fx;
        a = a + inc[3 + k] * fy;
        v[k] = a + inc[6 + k] * fz;
    }
}

/* backproject_cuda.cu:25-29 distance(p1, p2) with norm3df replaced by the IEEE expression. */
static inline float a1_distance(const float *p1, const float *p2)
{
    const float a = p1[0] - p2[0], b = p1[1] - p2[1], c = p1[2] - p2[2];
    const float s = (a * a + b * b) + c * c;
    return sqrtf(s);
}

/* backproject_cuda.cu:31-46 compute_distance then :155 (footprint 0, so the +-fp/2 term is
 * a subtraction of +0.0f, which changes no value).  Returns round((total-t0)/deltaT). */
static inline float a1_bin_of_distances(float lw, float cw, float d1, float d2, float t0, float dT)
{
    float total = lw + d1;                                 /* :45 left to right          */
    total = total + d2;
    total = total + cw;
    const float y = (total - t0) - 0.0f / 2.0f;            /* :155 voxel_footprint = 0   */
    const float q = y / dT;
    return roundf(q);
}

float or_a1_rounded_bin(const or_pair *pr, const float v[3], float t0, float dT)
{
    const float d1 = a1_distance(pr->laser_point, v);      /* :39 laser point -> voxel  */
    const float d2 = a1_distance(v, pr->camera_point);     /* :43 voxel -> camera point */
    return a1_bin_of_distances(pr->laser_wall, pr->camera_wall, d1, d2, t0, dT);
}

/* The same bin for a confocal pair given the squared distance s of its (single) wall point to
 * the voxel: d1 == d2 == sqrtf(s), since (a*a + b*b) + c*c is the same for v - w and w - v. */
float or_a1_confocal_bin(float lw, float cw, float s, float t0, float dT)
{
    const float d = sqrtf(s);
    return a1_bin_of_distances(lw, cw, d, d, t0, dT);
}

/* Threshold table of a confocal pair list: out[p][j] (j < n) = the smallest float s >= +0 with
 * or_a1_confocal_bin(lw_p, cw_p, s) >= j, found by plain bisection over the bit patterns of
 * the non-negative floats (ordered like their values; the bin never decreases in s).  +inf
 * satisfies every j (its bin is +inf). */
void or_confocal_thresholds(const float *lw, const float *cw, uint32_t P, uint32_t n, float t0, float dT,
                            float *out)
{
#pragma omp parallel for schedule(static)
    for (int64_t pj = 0; pj < (int64_t)P * n; ++pj) {
        const uint32_t p = (uint32_t)(pj / n), j = (uint32_t)(pj % n);
        const float jf = (float)j;
        uint32_t lo = 0u, hi = 0x7f800000u;   /* +0 .. +inf */
        float s;
        memcpy(&s, &lo, 4);
        if (or_a1_confocal_bin(lw[p], cw[p], s, t0, dT) >= jf) {
            out[pj] = s;
            continue;
        }
        while (hi - lo > 1u) {           /* invariant: bin(lo) < j <= bin(hi) */
            const uint32_t mid = lo + (hi - lo) / 2u;
            memcpy(&s, &mid, 4);
            if (or_a1_confocal_bin(lw[p], cw[p], s, t0, dT) >= jf) hi = mid;
            else lo = mid;
        }
        memcpy(&s, &hi, 4);
        out[pj] = s;
    }
}

/* The contributing bin index, or -1 when the single tap falls outside [0, T-1]
 * (backproject_cuda.cu:157-162 with min_time_index == max_time_index - 1). */
static inline int32_t a1_bin_index(float r, uint32_t T)
{
    if (r >= 0.0f && r <= (float)(T - 1)) return (int32_t)r;
    return -1;
}

/* Batch: bins for an explicit list of (voxel index, pair index) queries. */
void or_a1_bins(const or_pair *pairs, const uint32_t *voxel_xyz /* [n][3] */,
                const uint32_t *pair_idx, uint64_t n,
                const float zero[3], const float inc[9], float t0, float dT, uint32_t T,
                int32_t *out_bins, float *out_rounded)
{
    for (uint64_t i = 0; i < n; ++i) {
        float v[3];
        or_a1_voxel_position(zero, inc, voxel_xyz[3 * i], voxel_xyz[3 * i + 1],
                             voxel_xyz[3 * i + 2], v);
        const float r = or_a1_rounded_bin(&pairs[pair_idx[i]], v, t0, dT);
        if (out_rounded) out_rounded[i] = r;
        out_bins[i] = a1_bin_index(r, T);
    }
}

/* Volume in the reference summation order (backproject_cuda.cu:141-173): thread t of a
 * block of B threads sums pairs t, t+B, ..., for i < P/B (integer division, so the last
 * P mod B pairs are dropped, :150); thread 0 then adds partials 1..B-1 to its own (:169-171).
 * Output row-major [Nx][Ny][Nz], id = (x*Ny + y)*Nz + z (:82-84).  block == 0 selects the
 * reference launcher's B = min(256, P) (backproject_cuda.cu:418). */
void or_a1_volume_f32(const float *H, uint32_t T, const or_pair *pairs, uint32_t P,
                      const uint32_t N[3], const float zero[3], const float inc[9],
                      float t0, float dT, uint32_t block, float *out)
{
    uint32_t B = block ? block : (P < 256u ? P : 256u);
    if (B == 0) B = 1;
    const uint32_t iters = P / B;
    float *partial_all = NULL;
    const int64_t nvox = (int64_t)N[0] * N[1] * N[2];
#pragma omp parallel
    {
        float *partial = (float *)malloc(sizeof(float) * B);
#pragma omp for schedule(dynamic, 16)
        for (int64_t id = 0; id < nvox; ++id) {
            const uint32_t z = (uint32_t)(id % N[2]);
            const uint32_t y = (uint32_t)((id / N[2]) % N[1]);
            const uint32_t x = (uint32_t)(id / ((int64_t)N[2] * N[1]));
            float v[3];
            or_a1_voxel_position(zero, inc, x, y, z, v);
            for (uint32_t t = 0; t < B; ++t) {
                float acc = 0.0f;
                for (uint32_t i = 0; i < iters; ++i) {
                    const uint32_t p = i * B + t;
                    const int32_t b = a1_bin_index(or_a1_rounded_bin(&pairs[p], v, t0, dT), T);
                    if (b >= 0) acc = acc + H[(uint64_t)p * T + (uint32_t)b];
                }
                partial[t] = acc;
            }
            float s = partial[0];
            for (uint32_t t = 1; t < B; ++t) s = s + partial[t];
            out[id] = s;
        }
        free(partial);
    }
    (void)partial_all;
}

/* Same taps, every pair used, accumulated in double: the exact sum of the fp32 taps. */
void or_a1_volume_f64acc(const float *H, uint32_t T, const or_pair *pairs, uint32_t P,
                         const uint32_t N[3], const float zero[3], const float inc[9],
                         float t0, float dT, double *out)
{
    const int64_t nvox = (int64_t)N[0] * N[1] * N[2];
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t id = 0; id < nvox; ++id) {
        const uint32_t z = (uint32_t)(id % N[2]);
        const uint32_t y = (uint32_t)((id / N[2]) % N[1]);
        const uint32_t x = (uint32_t)(id / ((int64_t)N[2] * N[1]));
        float v[3];
        or_a1_voxel_position(zero, inc, x, y, z, v);
        double acc = 0.0;
        for (uint32_t p = 0; p < P; ++p) {
            const int32_t b = a1_bin_index(or_a1_rounded_bin(&pairs[p], v, t0, dT), T);
            if (b >= 0) acc += (double)H[(uint64_t)p * T + (uint32_t)b];
        }
        out[id] = acc;
    }
}

/* or_a1_volume_f64acc restricted to a list of voxels (for full-size captures). */
void or_a1_voxels_f64acc(const float *H, uint32_t T, const or_pair *pairs, uint32_t P,
                         const float zero[3], const float inc[9], float t0, float dT,
                         const uint32_t *voxel_xyz, int64_t n, double *out)
{
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t i = 0; i < n; ++i) {
        float v[3];
        or_a1_voxel_position(zero, inc, voxel_xyz[3 * i], voxel_xyz[3 * i + 1], voxel_xyz[3 * i + 2], v);
        double acc = 0.0;
        for (uint32_t p = 0; p < P; ++p) {
            const int32_t b = a1_bin_index(or_a1_rounded_bin(&pairs[p], v, t0, dT), T);
            if (b >= 0) acc += (double)H[(uint64_t)p * T + (uint32_t)b];
        }
        out[i] = acc;
    }
}

/* Complex variant (FT = cuComplex, backproject_cuda.cu:110-115 cuCaddf = componentwise
 * float adds), same order as or_a1_volume_f32.  H and out are interleaved (re, im). */
void or_a1_volume_c64(const float *H, uint32_t T, const or_pair *pairs, uint32_t P,
                      const uint32_t N[3], const float zero[3], const float inc[9],
                      float t0, float dT, uint32_t block, float *out)
{
    uint32_t B = block ? block : (P < 256u ? P : 256u);
    if (B == 0) B = 1;
    const uint32_t iters = P / B;
    const int64_t nvox = (int64_t)N[0] * N[1] * N[2];
#pragma omp parallel
    {
        float *partial = (float *)malloc(sizeof(float) * 2 * B);
#pragma omp for schedule(dynamic, 16)
        for (int64_t id = 0; id < nvox; ++id) {
            const uint32_t z = (uint32_t)(id % N[2]);
            const uint32_t y = (uint32_t)((id / N[2]) % N[1]);
            const uint32_t x = (uint32_t)(id / ((int64_t)N[2] * N[1]));
            float v[3];
            or_a1_voxel_position(zero, inc, x, y, z, v);
            for (uint32_t t = 0; t < B; ++t) {
                float re = 0.0f, im = 0.0f;
                for (uint32_t i = 0; i < iters; ++i) {
                    const uint32_t p = i * B + t;
                    const int32_t b = a1_bin_index(or_a1_rounded_bin(&pairs[p], v, t0, dT), T);
                    if (b >= 0) {
                        re = re + H[2 * ((uint64_t)p * T + (uint32_t)b)];
                        im = im + H[2 * ((uint64_t)p * T + (uint32_t)b) + 1];
                    }
                }
                partial[2 * t] = re;
                partial[2 * t + 1] = im;
            }
            float sr = partial[0], si = partial[1];
            for (uint32_t t = 1; t < B; ++t) { sr = sr + partial[2 * t]; si = si + partial[2 * t + 1]; }
            out[2 * id] = sr;
            out[2 * id + 1] = si;
        }
        free(partial);
    }
}

/* ------------------------------------------------------------------------------------ */
/* FP64 oracle: the A.1 geometry evaluated entirely in binary64 (positions, distances,    */
/* bin rounding, sums).  It is the error yardstick, not a parity target: its bins differ   */
/* from the fp32 formula where the fp32 path length lands within rounding of a half-bin.   */
/* ------------------------------------------------------------------------------------ */
static inline double f64_distance(const double *a, const float *b)
{
    const double x = a[0] - (double)b[0], y = a[1] - (double)b[1], z = a[2] - (double)b[2];
    return sqrt(x * x + y * y + z * z);
}

void or_fp64_volume(const float *H, uint32_t T, const or_pair *pairs, uint32_t P,
                    const uint32_t N[3], const float zero[3], const float inc[9],
                    double t0, double dT, int complex_samples, double *out)
{
    const int64_t nvox = (int64_t)N[0] * N[1] * N[2];
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t id = 0; id < nvox; ++id) {
        const uint32_t z = (uint32_t)(id % N[2]);
        const uint32_t y = (uint32_t)((id / N[2]) % N[1]);
        const uint32_t x = (uint32_t)(id / ((int64_t)N[2] * N[1]));
        double v[3];
        for (int k = 0; k < 3; ++k)
            v[k] = (double)zero[k] + (double)inc[k] * x + (double)inc[3 + k] * y +
                   (double)inc[6 + k] * z;
        double re = 0.0, im = 0.0;
        for (uint32_t p = 0; p < P; ++p) {
            const double total = (double)pairs[p].laser_wall + f64_distance(v, pairs[p].laser_point) +
                                 f64_distance(v, pairs[p].camera_point) + (double)pairs[p].camera_wall;
            const double r = round((total - t0) / dT);
            if (r >= 0.0 && r <= (double)(T - 1)) {
                const uint64_t i = (uint64_t)p * T + (uint64_t)r;
                if (complex_samples) { re += H[2 * i]; im += H[2 * i + 1]; }
                else re += H[i];
            }
        }
        if (complex_samples) { out[2 * id] = re; out[2 * id + 1] = im; }
        else out[id] = re;
    }
}

/* or_fp64_volume restricted to a list of voxels (real samples). */
void or_fp64_voxels(const float *H, uint32_t T, const or_pair *pairs, uint32_t P,
                    const float zero[3], const float inc[9], double t0, double dT,
                    const uint32_t *voxel_xyz, int64_t n, double *out)
{
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t i = 0; i < n; ++i) {
        const uint32_t x = voxel_xyz[3 * i], y = voxel_xyz[3 * i + 1], z = voxel_xyz[3 * i + 2];
        double v[3];
        for (int k = 0; k < 3; ++k)
            v[k] = (double)zero[k] + (double)inc[k] * x + (double)inc[3 + k] * y + (double)inc[6 + k] * z;
        double acc = 0.0;
        for (uint32_t p = 0; p < P; ++p) {
            const double total = (double)pairs[p].laser_wall + f64_distance(v, pairs[p].laser_point) +
                                 f64_distance(v, pairs[p].camera_point) + (double)pairs[p].camera_wall;
            const double r = round((total - t0) / dT);
            if (r >= 0.0 && r <= (double)(T - 1)) acc += H[(uint64_t)p * T + (uint64_t)r];
        }
        out[i] = acc;
    }
}

/* ------------------------------------------------------------------------------------ */
/* A.2 -- the CPU path formula (windowed)                                                */
/* ------------------------------------------------------------------------------------ */

/* backproject.hpp:42-53: float differences, squared and summed in double, double sqrt. */
static inline double a2_distance(const float *p1, const float *p2)
{
    double t[3];
    for (int i = 0; i < 3; ++i) {
        t[i] = (double)(p1[i] - p2[i]);
        t[i] = t[i] * t[i];
    }
    return sqrt(t[0] + t[1] + t[2]);
}

/* int conversion of a rounded float following max(0, (int)r) / min(T-1, (int)r) of
 * backproject.hpp:88-89 without the undefined out-of-range conversion. */
static inline int32_t a2_first(float r)
{
    if (!(r >= 0.0f)) return 0;                 /* negative or NaN -> 0 */
    if (r >= 2147483520.0f) return INT32_MAX;
    return (int32_t)r;
}
static inline int32_t a2_last(float r, uint32_t T)
{
    if (r != r || r <= -2147483520.0f) return INT32_MIN;  /* NaN: x86 converts to INT_MIN */
    if (r >= (float)(T - 1)) return (int32_t)T - 1;
    return (int32_t)r;
}

/* Grid of the CPU path: voxel_size = size / N (octree_volume.hpp:42, 157), corner
 * zero = position - size/2 (:149, :162), position = zero + (float)i * voxel_size (:164-166).
 * The reference keeps `zero` in a function-static variable, so a second call in the same
 * process silently reuses the first call's corner; this restatement uses each call's own
 * corner (documented deviation). */
void or_a2_grid(const float volume_position[3], const float volume_size[3], const uint32_t N[3],
                int diag_sum_in_double, float zero[3], float vsize[3], float *diagonal)
{
    for (int k = 0; k < 3; ++k) {
        zero[k] = volume_position[k] - volume_size[k] / 2.0f;
        vsize[k] = volume_size[k] / (float)N[k];
    }
    /* backproject.hpp:147: sqrt(sum(voxel_size * voxel_size)).  The squares are float;
     * xtensor's reducer promotes the float sum to double (version dependent: unpinned). */
    const float a = vsize[0] * vsize[0], b = vsize[1] * vsize[1], c = vsize[2] * vsize[2];
    if (diag_sum_in_double) *diagonal = (float)sqrt(((double)a + (double)b) + (double)c);
    else *diagonal = sqrtf((a + b) + c);
}

static inline void a2_window(const or_pair *pr, const float v[3], float t0, float dT,
                             float width, uint32_t T, int32_t *first, int32_t *last)
{
    /* backproject.hpp:85-87 */
    const float wvw = (float)(a2_distance(pr->laser_point, v) + a2_distance(v, pr->camera_point));
    float total = pr->laser_wall + wvw;
    total = total + pr->camera_wall;
    /* :88-89 */
    const float half = width / 2.0f;
    const float lo = roundf(((total - half) - t0) / dT);
    const float hi = roundf(((total + half) - t0) / dT);
    *first = a2_first(lo);
    *last = a2_last(hi, T);
}

/* backproject.hpp:138-216 classic_backprojection(depth = -1) + :75-96 accumulate_radiance:
 * per voxel, pairs in order, bins [first, last) in order, one float accumulator.  The voxel
 * loop is split into contiguous ranges over OpenMP threads as in :161-206 (this affects
 * only timing, not values).  The optional voxel range [vox_begin, vox_end) restricts work
 * to a subset (used to time a bounded CPU sample). */
void or_a2_volume(const float *H, uint32_t T, const or_pair *pairs, uint32_t P,
                  const uint32_t N[3], const float volume_position[3], const float volume_size[3],
                  float t0, float dT, int complex_samples, int diag_sum_in_double,
                  int64_t vox_begin, int64_t vox_end, float *out)
{
    float zero[3], vs[3], width;
    or_a2_grid(volume_position, volume_size, N, diag_sum_in_double, zero, vs, &width);
    const int64_t nvox = (int64_t)N[0] * N[1] * N[2];
    if (vox_end <= 0 || vox_end > nvox) vox_end = nvox;
    if (vox_begin < 0) vox_begin = 0;
#pragma omp parallel
    {
#ifdef _OPENMP
        const int64_t nth = omp_get_num_threads(), tid = omp_get_thread_num();
#else
        const int64_t nth = 1, tid = 0;
#endif
        const int64_t len = vox_end - vox_begin;
        const int64_t from = vox_begin + (len * tid) / nth;
        const int64_t to = vox_begin + (len * (tid + 1)) / nth;
        for (int64_t id = from; id < to; ++id) {
            const uint32_t z = (uint32_t)(id % N[2]);
            const uint32_t y = (uint32_t)((id / N[2]) % N[1]);
            const uint32_t x = (uint32_t)(id / ((int64_t)N[2] * N[1]));
            const float v[3] = {zero[0] + (float)x * vs[0], zero[1] + (float)y * vs[1],
                                zero[2] + (float)z * vs[2]};
            float re = 0.0f, im = 0.0f;
            for (uint32_t p = 0; p < P; ++p) {
                int32_t first, last;
                a2_window(&pairs[p], v, t0, dT, width, T, &first, &last);
                for (int32_t t = first; t < last; ++t) {
                    const uint64_t i = (uint64_t)p * T + (uint32_t)t;
                    if (complex_samples) { re = re + H[2 * i]; im = im + H[2 * i + 1]; }
                    else re = re + H[i];
                }
            }
            if (complex_samples) { out[2 * id] = re; out[2 * id + 1] = im; }
            else out[id] = re;
        }
    }
}

/* FP64 counterpart of A.2 (window edges and sums in double) -- error yardstick only. */
void or_a2_volume_fp64(const float *H, uint32_t T, const or_pair *pairs, uint32_t P,
                       const uint32_t N[3], const float volume_position[3],
                       const float volume_size[3], double t0, double dT, double *out)
{
    double zero[3], vs[3];
    for (int k = 0; k < 3; ++k) {
        zero[k] = (double)volume_position[k] - (double)volume_size[k] / 2.0;
        vs[k] = (double)volume_size[k] / (double)N[k];
    }
    const double width = sqrt(vs[0] * vs[0] + vs[1] * vs[1] + vs[2] * vs[2]);
    const int64_t nvox = (int64_t)N[0] * N[1] * N[2];
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t id = 0; id < nvox; ++id) {
        const uint32_t z = (uint32_t)(id % N[2]);
        const uint32_t y = (uint32_t)((id / N[2]) % N[1]);
        const uint32_t x = (uint32_t)(id / ((int64_t)N[2] * N[1]));
        const double v[3] = {zero[0] + x * vs[0], zero[1] + y * vs[1], zero[2] + z * vs[2]};
        double acc = 0.0;
        for (uint32_t p = 0; p < P; ++p) {
            const double total = (double)pairs[p].laser_wall + f64_distance(v, pairs[p].laser_point) +
                                 f64_distance(v, pairs[p].camera_point) + (double)pairs[p].camera_wall;
            double lo = round((total - width / 2 - t0) / dT), hi = round((total + width / 2 - t0) / dT);
            if (lo < 0) lo = 0;
            if (hi > (double)T - 1) hi = (double)T - 1;
            for (int64_t t = (int64_t)lo; t < (int64_t)hi; ++t) acc += H[(uint64_t)p * T + (uint64_t)t];
        }
        out[id] = acc;
    }
}

/* ------------------------------------------------------------------------------------ */
/* Host pipeline (backproject.hpp:356-418, 490-740, 820-847)                             */
/* ------------------------------------------------------------------------------------ */

/* backproject.hpp:42-53 distance, returned as double and stored by the caller as float. */
static inline float wall_distance(const float *origin, const float *point)
{
    return (float)a2_distance(origin, point);
}

/* precompute_distances, backproject.hpp:356-418.  Grids are row-major [X][Y][3].
 * Exhaustive: p = ((lx*Ly + ly)*Cx + cx)*Cy + cy (:386-389).
 * Confocal: p = lx*Ly + ly, BOTH legs use the camera grid point and camera_position
 * (:405-411); laser grid values and laser_position are ignored. */
void or_precompute_pairs(const float laser_position[3], const float *laser_grid, uint32_t Lx,
                         uint32_t Ly, const float camera_position[3], const float *camera_grid,
                         uint32_t Cx, uint32_t Cy, int confocal, or_pair *out)
{
    if (confocal) {
        for (uint32_t lx = 0; lx < Lx; ++lx)
            for (uint32_t ly = 0; ly < Ly; ++ly) {
                const float *cp = camera_grid + 3 * ((uint64_t)lx * Cy + ly);
                or_pair *pr = &out[(uint64_t)lx * Ly + ly];
                memcpy(pr->laser_point, cp, sizeof(float) * 3);
                memcpy(pr->camera_point, cp, sizeof(float) * 3);
                const float w = wall_distance(camera_position, cp);
                pr->camera_wall = w;
                pr->laser_wall = w;
            }
        return;
    }
    for (uint32_t lx = 0; lx < Lx; ++lx)
        for (uint32_t ly = 0; ly < Ly; ++ly)
            for (uint32_t cx = 0; cx < Cx; ++cx)
                for (uint32_t cy = 0; cy < Cy; ++cy) {
                    const uint64_t p = (((uint64_t)lx * Ly + ly) * Cx + cx) * Cy + cy;
                    const float *lp = laser_grid + 3 * ((uint64_t)lx * Ly + ly);
                    const float *cp = camera_grid + 3 * ((uint64_t)cx * Cy + cy);
                    memcpy(out[p].laser_point, lp, sizeof(float) * 3);
                    memcpy(out[p].camera_point, cp, sizeof(float) * 3);
                    out[p].camera_wall = wall_distance(camera_position, cp);
                    out[p].laser_wall = wall_distance(laser_position, lp);
                }
}

/* Voxel basis, backproject.hpp:820-839: zero = position - size/2 and
 * inc[r][c] = axis_r[c] * size[c] / (N[c] - 1).  A non-zero wall_normal builds
 * y = -n, x = y cross z, z = y cross x, each normalised (:825-833); the norm's sum is
 * taken in double (xtensor reducer promotion, unpinned). */
void or_basis(const float volume_position[3], const float volume_size[3], const uint32_t N[3],
              const float wall_normal[3], float zero[3], float inc[9])
{
    float ax[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    if (wall_normal[0] != 0.0f || wall_normal[1] != 0.0f || wall_normal[2] != 0.0f) {
        float y[3] = {-wall_normal[0], -wall_normal[1], -wall_normal[2]};
        float z[3] = {0, 0, 1};
        float x[3] = {y[1] * z[2] - y[2] * z[1], y[2] * z[0] - y[0] * z[2], y[0] * z[1] - y[1] * z[0]};
        float z2[3] = {y[1] * x[2] - y[2] * x[1], y[2] * x[0] - y[0] * x[2], y[0] * x[1] - y[1] * x[0]};
        float *v3[3] = {x, y, z2};
        for (int r = 0; r < 3; ++r) {
            const float *a = v3[r];
            const float aa = a[0] * a[0], bb = a[1] * a[1], cc = a[2] * a[2];
            const double nrm = sqrt(((double)aa + (double)bb) + (double)cc);
            for (int c = 0; c < 3; ++c) ax[r][c] = (float)((double)a[c] / nrm);
        }
    }
    for (int k = 0; k < 3; ++k) zero[k] = volume_position[k] - volume_size[k] / 2.0f;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            inc[3 * r + c] = (ax[r][c] * volume_size[c]) / (float)(N[c] - 1u);
}

/* get_min_max_distances (backproject.hpp:579-598) over point sets of `dim`-vectors whose
 * first three entries are used by distance(). */
static void min_max_dist(const float *a, int na, int stride_a, const float *b, int nb,
                         int stride_b, float *mn, float *mx)
{
    float lo = FLT_MAX, hi = 0.0f;
    for (int i = 0; i < na; ++i)
        for (int j = 0; j < nb; ++j) {
            const float d = (float)a2_distance(a + (size_t)i * stride_a, b + (size_t)j * stride_b);
            lo = d < lo ? d : lo;
            hi = d > hi ? d : hi;
        }
    *mn = lo;
    *mx = hi;
}

/* get_time_limits, backproject.hpp:617-674, row-major inputs.
 * - Laser grid corners: grid[0,0], grid[0,Y-1], grid[X-1,0], grid[X-1,Y-1] (:490-547).
 * - Camera grid: for row-major data the reference transposes the camera grid to [3][Y][X]
 *   before taking corners (:651-652), so each "corner" is a length-X row of that array and
 *   distance() reads its first three entries: (G[0][y][c], G[1][y][c], G[2][y][c]) for
 *   c in {0, 2} and y in {0, Y-1}.  Restated as such (needs X >= 3).
 * - Volume corners: (pos - size/2) + mask * size (:514-525, 556-569).
 * - Origins: the reference indexes the 1-D origin array with two slices (:592, 660-661),
 *   which is undefined inside xtensor; restated as the single origin point.
 * Returns min_T_index - 100 and max_T_index + 100 as floats (:665-673). */
int or_time_limits(const float laser_position[3], const float *laser_grid, uint32_t Lx,
                   uint32_t Ly, const float camera_position[3], const float *camera_grid,
                   uint32_t Cx, uint32_t Cy, const float volume_size[3],
                   const float volume_position[3], float t0, float dT, float *min_idx,
                   float *max_idx)
{
    if (Cx < 3) return -1;
    float las[4][3], cam[4][3], vol[8][3];
    const uint32_t lxs[2] = {0, Lx - 1}, lys[2] = {0, Ly - 1};
    for (int m = 0; m < 4; ++m) {
        const uint32_t x = lxs[(m >> 1) & 1], y = lys[m & 1];
        memcpy(las[m], laser_grid + 3 * ((uint64_t)x * Ly + y), sizeof(float) * 3);
    }
    /* transposed camera grid T[c][y][x] = G[x][y][c]; corners over (c in {0,2}, y in {0,Y-1}) */
    const uint32_t cs[2] = {0, 2}, cys[2] = {0, Cy - 1};
    for (int m = 0; m < 4; ++m) {
        const uint32_t c = cs[(m >> 1) & 1], y = cys[m & 1];
        for (int k = 0; k < 3; ++k) cam[m][k] = camera_grid[3 * ((uint64_t)k * Cy + y) + c];
    }
    for (int m = 0; m < 8; ++m)
        for (int k = 0; k < 3; ++k) {
            const float lower = volume_position[k] - volume_size[k] / 2.0f;
            const float mult = (m & (1 << (2 - k))) ? 1.0f : 0.0f;
            vol[m][k] = lower + mult * volume_size[k];
        }
    float mn_lw, mx_lw, mn_cw, mx_cw, mn_lv, mx_lv, mn_cv, mx_cv;
    min_max_dist(laser_position, 1, 3, &las[0][0], 4, 3, &mn_lw, &mx_lw);
    min_max_dist(camera_position, 1, 3, &cam[0][0], 4, 3, &mn_cw, &mx_cw);
    min_max_dist(&vol[0][0], 8, 3, &las[0][0], 4, 3, &mn_lv, &mx_lv);
    min_max_dist(&vol[0][0], 8, 3, &cam[0][0], 4, 3, &mn_cv, &mx_cv);
    float mn = mn_lw + mn_cw;
    mn = mn + mn_lv;
    mn = mn + mn_cv;
    float mx = mx_lw + mx_cw;
    mx = mx + mx_lv;
    mx = mx + mx_cv;
    mn = mn - t0;
    mx = mx - t0;
    *min_idx = mn / dT - 100.0f;
    *max_idx = mx / dT + 100.0f;
    return 0;
}

/* get_transient_chunk, backproject.hpp:689-740 (row-major): rows of length T copied into a
 * zero-padded chunk covering [floor(min), ceil(max)).  Returns the chunk length Tc. */
uint32_t or_chunk_length(float min_idx, float max_idx)
{
    return (uint32_t)(ceilf(max_idx) - floorf(min_idx));
}

void or_transient_chunk(const float *data, uint64_t rows, uint32_t T, float min_idx,
                        float max_idx, int complex_samples, float *chunk)
{
    const float fmin = floorf(min_idx), fmax = ceilf(max_idx);
    const uint64_t Tc = (uint64_t)(fmax - fmin);
    const uint64_t orig_from = fmin > 0.0f ? (uint64_t)fmin : 0;
    uint64_t orig_to = (uint64_t)(fmax > 0.0f ? fmax : 0.0f);
    if (orig_to > T) orig_to = T;
    const uint64_t chunk_from = fmin > 0.0f ? 0 : (uint64_t)fabsf(fmin);
    const int c = complex_samples ? 2 : 1;
    memset(chunk, 0, sizeof(float) * c * rows * Tc);
    if (orig_to <= orig_from) return;
    const uint64_t n = orig_to - orig_from;
    for (uint64_t r = 0; r < rows; ++r) {
        if (chunk_from + n > Tc) continue;
        memcpy(chunk + c * (r * Tc + chunk_from), data + c * (r * T + orig_from),
               sizeof(float) * c * n);
    }
}

/* chunked_t0 = t0 + min_T_index * deltaT with the unfloored min (backproject.hpp:847). */
float or_chunked_t0(float t0, float min_idx, float dT)
{
    return t0 + min_idx * dT;
}

/* ------------------------------------------------------------------------------------ */
/* Phasor-field filter, backproject.hpp:430-479                                           */
/* ------------------------------------------------------------------------------------ */

/* Taps: pulse_size = (int)(times*wl/dT) and cycle_size = (int)(wl/dT) in float (:434-436);
 * sigma = times*wl/6 (:439); t_k = dT*(k - pulse_size/2.0) for k = 1..pulse_size in double
 * (:442); g_k = exp(-t^2/(2*sigma*sigma)) with the denominator in float (:443);
 * progression = 2*pi*((1.0/cycle_size)*k) in double (:445); sin/cos stored as float
 * (:447-448); taps = float(float(cos) * g) (:450-451).  Returns pulse_size. */
int or_phasor_taps(float wavelength, float dT, uint32_t times, float *cos_taps, float *sin_taps,
                   int max_taps)
{
    const int pulse = (int)(((float)times * wavelength) / dT);
    const int cycle = (int)(wavelength / dT);
    if (pulse <= 0 || cycle <= 0 || pulse > max_taps) return -pulse - 1;
    const float vsigma = ((float)times * wavelength) / 6.0f;
    const float denom = (2.0f * vsigma) * vsigma;
    for (int k = 1; k <= pulse; ++k) {
        const double t = (double)dT * ((double)k - pulse / 2.0);
        const double g = exp(-t * t / (double)denom);
        const double prog = 2.0 * M_PI * ((1.0 / cycle) * (double)k);
        const float s = (float)sin(prog), c = (float)cos(prog);
        cos_taps[k - 1] = (float)((double)c * g);
        sin_taps[k - 1] = (float)((double)s * g);
    }
    return pulse;
}

/* out[t] = sum_{j<pulse} in[t + j - pulse/2] * (cos_j + i sin_j) over 0 < idx < Tc
 * (:459-476; bin 0 is never read, :468), complex<float> accumulation in j order. */
void or_phasor_filter(const float *in, uint64_t rows, uint32_t Tc, const float *cos_taps,
                      const float *sin_taps, int pulse, float *out)
{
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < (int64_t)rows; ++r)
        for (int64_t t = 0; t < (int64_t)Tc; ++t) {
            float re = 0.0f, im = 0.0f;
            for (int j = 0; j < pulse; ++j) {
                const int64_t idx = t + j - pulse / 2;
                if (idx > 0 && idx < (int64_t)Tc) {
                    const float x = in[r * Tc + idx];
                    re = re + x * cos_taps[j];
                    im = im + x * sin_taps[j];
                }
            }
            out[2 * (r * Tc + t)] = re;
            out[2 * (r * Tc + t) + 1] = im;
        }
}

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
