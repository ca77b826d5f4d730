/*
 * bpx.h -- C ABI of the MI355X-native NLOS time-of-flight backprojector (libbpx.so).
 *
 * Plain C: POD structs, pointers and sizes, status codes.  No function calls exit(); every
 * failure returns a bpx_status and leaves a message for bpx_last_error().  Host-buffer entry
 * points are synchronous; device-buffer entry points (plans, batches) enqueue on the stream
 * given (NULL = the library's own stream) and synchronise only where documented.  Executions
 * of one plan or batch share its buffers, so each is ordered (stream-side, no host wait) after
 * the previous one, whatever stream that used; the synchronize, read, stats and destroy calls
 * wait for the last execution on any stream.
 *
 * Reference interfaces replaced (mjgalindo/backprojector):
 *   bpx_backprojection          <- call_cuda_backprojection          backproject_cuda.hpp:19-28
 *   bpx_octree_backprojection   <- call_cuda_octree_backprojection   backproject_cuda.hpp:30-39
 *   bpx_complex_backprojection  <- call_cuda_complex_backprojection  backproject_cuda.hpp:41-49
 *   bpx_window_backprojection   <- classic_backprojection (CPU path) backproject.hpp:236-247
 *   bpx_phasor_pulse            <- phasor_pulse                      backproject.hpp:430-479
 *   bpx_nlos_backproject        <- bp::backproject via nlosbpy.backproject
 *                                  backproject.hpp:762-909, nlosbpy.cpp:15-45
 *   bpx_nlos_phasor             <- bp::phasor_reconstruction via nlosbpy.phasor
 *                                  backproject.hpp:932-1039, nlosbpy.cpp:47-77
 *   bpx_pair                    <- struct CameraLaserPair            backproject_cuda.hpp:9-13
 */
#ifndef BPX_H
#define BPX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BPX_ABI_VERSION 1

typedef enum bpx_status {
    BPX_OK = 0,
    BPX_ERR_INVALID_ARGUMENT = 1, /* shapes, sizes, non-finite geometry, bad enum values */
    BPX_ERR_OUT_OF_MEMORY = 2,    /* device or host allocation failed */
    BPX_ERR_DEVICE = 3,           /* a HIP call failed (message names it) */
    BPX_ERR_UNSUPPORTED = 4,      /* valid request outside what this library implements */
    BPX_ERR_NO_DEVICE = 5         /* no gfx950 device visible */
} bpx_status;

/* Same layout and meaning as the reference's CameraLaserPair (32 bytes). */
typedef struct bpx_pair {
    float camera_point[3];
    float laser_point[3];
    float camera_wall;
    float laser_wall;
} bpx_pair;

/* Device storage of transient samples. Real captures: F32/F16/BF16; complex (phasor)
 * captures are stored as the complex counterpart (c64 / 2xf16 / 2xbf16). 16-bit storage
 * uses a per-capture power-of-two scale chosen from max|H| (exact rescaling). */
typedef enum bpx_storage {
    BPX_STORAGE_F32 = 0,
    BPX_STORAGE_F16 = 1,
    BPX_STORAGE_BF16 = 2
} bpx_storage;

/* Semantics of a reconstruction.
 * SINGLE_TAP: the reference GPU path (backproject_cuda.cu:117-175): one bin per pair,
 *   grid spacing size/(N-1) with the wall-normal basis, fp32 path length.
 * VOXEL_WINDOW: the reference CPU path (backproject.hpp:75-96, 138-216): bins in the
 *   voxel-diagonal window, grid spacing size/N, double-precision distances. */
typedef enum bpx_semantics {
    BPX_SEMANTICS_SINGLE_TAP = 0,
    BPX_SEMANTICS_VOXEL_WINDOW = 1
} bpx_semantics;

/* Options of the extended entry points; a NULL pointer means all defaults. */
typedef struct bpx_options {
    uint32_t storage;     /* bpx_storage (default F32)                                       */
    uint32_t kahan_block; /* pairs summed plainly between compensated adds (0 -> 32)          */
    uint32_t flags;       /* BPX_FLAG_*                                                       */
    int32_t variant;      /* gather-kernel variant: -1 default (the LDS-staged kernel where it
                             applies: fp32 real samples, separable grid, flat wall, short bin
                             windows; else 5); 0-2 split tables, 3-5 fused rows with 4, 8, 2
                             voxels per lane; 6 / 7 / 8 LDS-staged 16^3 tiles of 256 lanes,
                             16x16x8 of 256, 16^3 of 512                                    */
} bpx_options;

/* Drop the last P mod min(256, P) pairs exactly as the reference launcher does
 * (backproject_cuda.cu:150, 418); off by default (every pair is used). */
#define BPX_FLAG_REFERENCE_PAIR_TRUNCATION 0x1u
/* bpx_nlos_plan_create with a non-zero wall normal: build the volume basis right-handed and
 * centred on volume_position (axis r spaced volume_size[r] / (N[r] - 1); y = -normal, x from
 * y x world-z, or y x world-x when the normal is within ~25 degrees of z), instead of the
 * reference's construction (backproject.hpp:825-839), which gives a negative z axis for the
 * common normal (0, -1, 0) (README.md:116) and NaN axes for a normal parallel to z.  No effect
 * when the normal is zero; the identity basis results for (0, -1, 0). */
#define BPX_FLAG_CENTRED_WALL_BASIS 0x2u
/* bpx_batch_create: every capture's transient_chunk holds interleaved complex samples
 * (transient_size counts complex elements) and every volume is complex (2 floats per voxel). */
#define BPX_FLAG_COMPLEX_SAMPLES 0x4u
/* F32 storage on the LDS-staged confocal kernel: when every sample of an execution is an
 * integer of magnitude <= 2048 (photon counts, as captures hold), the rows are staged as fp16,
 * which holds those values exactly, and the gather reads half the LDS bytes per lookup; the
 * volume is the fp32 rows' bit for bit.  Each execution checks its samples on the device.
 * This flag keeps fp32 rows always (no reference counterpart: an internal layout choice). */
#define BPX_FLAG_KEEP_F32_ROWS 0x8u

/* ------------------------------------------------------------------ diagnostics */
const char* bpx_last_error(void);   /* thread-local; "" when the last call succeeded */
const char* bpx_version(void);
int bpx_abi_version(void);
/* Number of usable gfx950 devices (0 when none); never fails. */
int bpx_device_count(void);
/* The per-pair element windows [lo, hi) of the fused rows that a confocal plan of this
 * geometry stages at each execution: every element a gather kernel can read for the pair and
 * any voxel of the grid (element j of a row holds bin j - 1's sample and bin j's threshold).
 * No reference counterpart: exposed so that the windows can be checked against the bins on
 * the host.  pairs: confocal (laser point == camera point); Tc: bins of the chunk; out:
 * 2 * num_pairs uint32 (lo, hi), within [0, roundup(Tc + 2, 16)].  Host only (no device call). */
bpx_status bpx_row_windows(const bpx_pair* pairs, uint32_t num_pairs, const uint32_t voxels_per_side[3],
                           const float volume_zero_pos[3], const float voxel_inc[9], float t0, float deltaT,
                           uint32_t Tc, uint32_t* out);

/* ------------------------------------------------------------------ reference-equivalent
 * Host buffers in, host volume out (overwritten), synchronous.
 * transient_chunk: [num_pairs][T] row-major samples, transient_size = num_pairs * T.
 * voxel_volume: [N0][N1][N2] row-major; voxel_inc: 3x3 row-major (row r = step of index r). */
bpx_status bpx_backprojection(const float* transient_chunk, uint64_t transient_size, uint32_t T,
                              const bpx_pair* scanned_pairs, uint32_t num_pairs,
                              float* voxel_volume, const uint32_t voxels_per_side[3],
                              const float volume_zero_pos[3], const float voxel_inc[9],
                              float t0, float deltaT);

/* Complex samples and volume as interleaved (re, im) float pairs; transient_size counts
 * complex elements. */
bpx_status bpx_complex_backprojection(const float* transient_chunk, uint64_t transient_size,
                                      uint32_t T, const bpx_pair* scanned_pairs,
                                      uint32_t num_pairs, float* voxel_volume,
                                      const uint32_t voxels_per_side[3],
                                      const float volume_zero_pos[3], const float voxel_inc[9],
                                      float t0, float deltaT);

bpx_status bpx_backprojection_ex(const float* transient_chunk, uint64_t transient_size,
                                 uint32_t T, const bpx_pair* scanned_pairs, uint32_t num_pairs,
                                 float* voxel_volume, const uint32_t voxels_per_side[3],
                                 const float volume_zero_pos[3], const float voxel_inc[9],
                                 float t0, float deltaT, int complex_samples,
                                 const bpx_options* options);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the volume of
 * bpx_backprojection (fp32 samples, every pair used) computed on caller-owned DEVICE buffers,
 * with no hipMalloc / hipFree / allocating library call.  The library's working tables live in
 * one fixed device array of BPX_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends.
 *   transient_chunk: device pointer, [num_pairs][T] fp32 samples (transient_size = num_pairs * T)
 *   scanned_pairs:   HOST array (copied into the fixed array; finite values only)
 *   voxel_volume:    device pointer, [N0][N1][N2] fp32, overwritten
 * Work runs on hip_stream (NULL = the null stream); the call returns when it has finished.
 * Limits (else BPX_ERR_UNSUPPORTED): roundup(num_pairs * 32, 256) + (T + 1) * 4 <=
 * BPX_NOALLOC_SCRATCH_BYTES, T <= 16383.  Every call on one GPU uses that GPU's single array:
 * calls that overlap on one GPU (other threads, other streams) race. */
#define BPX_NOALLOC_SCRATCH_BYTES (8u << 20)
bpx_status bpx_backprojection_noalloc(const float* transient_chunk, uint64_t transient_size, uint32_t T,
                                      const bpx_pair* scanned_pairs, uint32_t num_pairs,
                                      float* voxel_volume, const uint32_t voxels_per_side[3],
                                      const float volume_zero_pos[3], const float voxel_inc[9],
                                      float t0, float deltaT, void* hip_stream);

/* CPU-path semantics on the GPU: the grid is given as the reference CPU path takes it
 * (volume_position = centre, volume_size), spacing size/N, voxel-diagonal window. */
bpx_status bpx_window_backprojection(const float* transient_chunk, uint64_t transient_size,
                                     uint32_t T, const bpx_pair* scanned_pairs,
                                     uint32_t num_pairs, float* voxel_volume,
                                     const uint32_t voxels_per_side[3],
                                     const float volume_position[3], const float volume_size[3],
                                     float t0, float deltaT, int complex_samples);

/* Phasor-field pulse filter.  in: [rows][T] real; out: [rows][T] interleaved complex.
 * sigma <= 0 selects the reference's sigma = times*wavelength/6 and pulse length
 * (int)(times*wavelength/deltaT); sigma > 0 is an extension: Gaussian window of that sigma
 * truncated at +-3 sigma (pulse = 2*ceil(3 sigma/deltaT) + 1 taps). */
bpx_status bpx_phasor_pulse(const float* transient_in, uint64_t rows, uint32_t T,
                            float wavelength, float deltaT, uint32_t times, float sigma,
                            float* complex_out);

/* ------------------------------------------------------------------ reference pipeline
 * Row-major inputs as nlosbpy passes them with assume_row_major=True:
 *   confocal: transient [X][Y][T], grids [X][Y][3]; exhaustive: transient [Lx][Ly][Cx][Cy][T].
 * shape arrays give the dimensions of each array (ndim entries).  assume_row_major = 0 means
 * the arrays are the column-major (Fortran-order) transposes; they are transposed to row-major
 * first.  Output: float [N0][N1][N2] (backproject) or interleaved complex (phasor). */
bpx_status bpx_nlos_backproject(const float* transient_data, const int64_t* data_shape, int data_ndim,
                                const float* camera_grid_positions, const int64_t* camera_grid_shape,
                                const float* laser_grid_positions, const int64_t* laser_grid_shape,
                                const float camera_position[3], const float laser_position[3],
                                float t0, float deltaT, int is_confocal,
                                const float volume_position[3], const float volume_size[3],
                                const uint32_t voxels_per_side[3], const float wall_normal[3],
                                int assume_row_major, int use_cpu, float* volume_out);

bpx_status bpx_nlos_phasor(const float* transient_data, const int64_t* data_shape, int data_ndim,
                           const float* camera_grid_positions, const int64_t* camera_grid_shape,
                           const float* laser_grid_positions, const int64_t* laser_grid_shape,
                           const float camera_position[3], const float laser_position[3],
                           float t0, float deltaT, int is_confocal,
                           const float volume_position[3], const float volume_size[3],
                           const uint32_t voxels_per_side[3], float wavelength,
                           int assume_row_major, int use_cpu, float* complex_volume_out);

/* bpx_nlos_phasor at several wavelengths of one capture (no reference counterpart; equal to
 * `count` bpx_nlos_phasor calls, volume i at complex_volumes_out + 2 * N0*N1*N2 * i): the
 * capture is copied and cropped once, and one plan is re-staged per wavelength. */
bpx_status bpx_nlos_phasor_multi(const float* transient_data, const int64_t* data_shape, int data_ndim,
                                 const float* camera_grid_positions, const int64_t* camera_grid_shape,
                                 const float* laser_grid_positions, const int64_t* laser_grid_shape,
                                 const float camera_position[3], const float laser_position[3],
                                 float t0, float deltaT, int is_confocal,
                                 const float volume_position[3], const float volume_size[3],
                                 const uint32_t voxels_per_side[3], const float* wavelengths, uint32_t count,
                                 int assume_row_major, int use_cpu, float* complex_volumes_out);

/* Output of bpx_nlos_phasor_output: the complex volume, or what the reference's callers derive
 * from it on the host -- the CLI writes std::abs of every voxel (main.cpp:162-166). */
typedef enum bpx_phasor_output {
    BPX_PHASOR_COMPLEX = 0,         /* interleaved complex, 2 * N0*N1*N2 floats per wavelength  */
    BPX_PHASOR_MAGNITUDE = 1,       /* |v|, N0*N1*N2 floats per wavelength                       */
    BPX_PHASOR_PHASE = 2,           /* atan2(im, re) in [-pi, pi], N0*N1*N2 floats per wavelength */
    BPX_PHASOR_MAGNITUDE_PHASE = 3  /* per wavelength: the magnitude volume, then the phase one  */
} bpx_phasor_output;

/* bpx_nlos_phasor_multi whose volumes leave the device as `output` says (no reference
 * counterpart: the reference returns the complex volume, backproject.hpp:932-1039, and its
 * CLI takes the magnitude on the host, main.cpp:162-166).  Magnitude = sqrt(re^2 + im^2)
 * evaluated in double and rounded once to float; phase = atan2(im, re) in double rounded to
 * float.  volumes_out holds count x (2, 1, 1 or 2) x N0*N1*N2 floats. */
bpx_status bpx_nlos_phasor_output(const float* transient_data, const int64_t* data_shape, int data_ndim,
                                  const float* camera_grid_positions, const int64_t* camera_grid_shape,
                                  const float* laser_grid_positions, const int64_t* laser_grid_shape,
                                  const float camera_position[3], const float laser_position[3],
                                  float t0, float deltaT, int is_confocal,
                                  const float volume_position[3], const float volume_size[3],
                                  const uint32_t voxels_per_side[3], const float* wavelengths, uint32_t count,
                                  int assume_row_major, int use_cpu, int output, float* volumes_out);

/* Intermediates of the reference pipeline up to the device call (backproject.hpp:779-847):
 * the transient chunk [rows][Tc], the pair table, the GPU-path basis and chunked t0.  Arrays
 * are allocated by the library; release them with bpx_nlos_prepared_free. */
typedef struct bpx_nlos_prepared {
    float* chunk;
    uint64_t rows;
    uint32_t Tc;
    bpx_pair* pairs;
    uint32_t num_pairs;
    float zero[3];
    float inc[9];
    float chunked_t0;
    float min_index, max_index;
} bpx_nlos_prepared;
bpx_status bpx_nlos_prepare(const float* transient_data, const int64_t* data_shape, int data_ndim,
                            const float* camera_grid_positions, const int64_t* camera_grid_shape,
                            const float* laser_grid_positions, const int64_t* laser_grid_shape,
                            const float camera_position[3], const float laser_position[3],
                            float t0, float deltaT, int is_confocal,
                            const float volume_position[3], const float volume_size[3],
                            const uint32_t voxels_per_side[3], const float wall_normal[3],
                            int assume_row_major, bpx_nlos_prepared* out);
void bpx_nlos_prepared_free(bpx_nlos_prepared* p);

/* ------------------------------------------------------------------ octree refinement
 * Coarse-to-fine reconstruction of the single-tap GPU path for confocal captures (the
 * reference's octree, backproject_cuda.cu:183-367, 503-669, is non-functional and prints
 * "GPU OCTREE BP NOT WORKING YET"; this is a new design).  The fine grid N (each side
 * divisible by 2^levels) is reached from N / 2^levels by `levels` refinements.  Per level,
 * every active block gets a sound upper bound U of its fine voxels' values (window maximum
 * of each pair's samples over the block's bin window, summed over pairs) and the exact value
 * of one representative fine voxel; the running maximum and the level's L2 norm are reduced
 * in FP64.  A block is refined iff U >= threshold * running max (and, with a learned
 * proposer attached, iff its proposal probability exceeds proposal_threshold).  Finest-level
 * voxels are computed exactly (bit-identical taps to bpx_backprojection); voxels never
 * reached are 0.  With samples >= 0 every voxel whose value is >= threshold * max is
 * computed exactly. */
typedef struct bpx_octree_options {
    uint32_t levels;          /* refinements (default 4: 16^3 -> 256^3)                  */
    float threshold;          /* tau (default 0.25)                                      */
    float proposal_threshold; /* learned proposals: refine iff probability > this (0.5) */
    uint32_t flags;           /* BPX_OCTREE_*                                            */
    float level_proposal_threshold[8]; /* per coarse level (0 = use proposal_threshold)  */
    float keep_bound_ratio;   /* learned proposals: a child whose block bound U reaches
                                 keep_bound_ratio x the running max is refined whatever the
                                 net says (the sound rule as the fallback where the bound is
                                 high); 0 = the net alone decides                           */
    float level_fallback_threshold[8]; /* learned proposals, per coarse level: a child the
                                 net does not refine but whose probability exceeds this is
                                 computed at full resolution instead (every fine voxel of it,
                                 exactly: the plain reconstruction as the low-confidence
                                 fallback); 0 = no fallback band                            */
    uint32_t coarse_pair_stride; /* learned proposals: the coarse levels' features (U, V) use
                                 every s-th wall point along each grid axis (s x s fewer
                                 pairs; 1 = all pairs); finest-level voxels always use every
                                 pair.  Used while a net is attached and for bounds-only
                                 executions (training data); the bound rule alone always uses
                                 every pair, so that it stays sound                          */
} bpx_octree_options;

/* Run the coarse levels only and skip the exact finest level (training data for the learned
 * proposals; with threshold 0 every block of every coarse level is evaluated). */
#define BPX_OCTREE_BOUNDS_ONLY 0x1u

typedef struct bpx_octree_stats {
    uint32_t levels;
    uint32_t active[9];       /* blocks evaluated at levels 0..levels-1, fine voxels last */
    double running_max[8];    /* FP64 running max of exact representative values        */
    double bound_max[8];      /* FP64 max of the level's upper bounds                   */
    double bound_norm[8];     /* FP64 L2 norm of the level's upper bounds               */
    double threshold[8];      /* threshold applied at each level                        */
    uint64_t fine_updates;    /* voxel-pair updates of the exact finest level           */
    uint64_t bound_updates;   /* block-pair bound evaluations of all coarse levels       */
    uint32_t fallback_voxels; /* fine voxels computed by the full-resolution fallback
                                 (included in active[levels])                             */
    uint32_t coarse_pairs;    /* pairs the coarse levels used                            */
} bpx_octree_stats;

/* Same arguments as call_cuda_octree_backprojection (backproject_cuda.hpp:30-39). */
bpx_status bpx_octree_backprojection(const float* transient_chunk, uint64_t transient_size,
                                     uint32_t T, const bpx_pair* scanned_pairs, uint32_t num_pairs,
                                     float* voxel_volume, const uint32_t voxels_per_side[3],
                                     const float volume_zero_pos[3], const float voxel_inc[9],
                                     float t0, float deltaT, const bpx_octree_options* options,
                                     bpx_octree_stats* stats);

/* ------------------------------------------------------------------ device-resident plans
 * A plan owns the device copy of one capture (samples staged once into the padded layout
 * and storage format) and its output volume; bpx_plan_execute runs the whole per-capture
 * reconstruction (threshold tables + gather) on the device. */
typedef struct bpx_plan_s* bpx_plan;

bpx_status bpx_plan_create(const float* transient_chunk, uint64_t transient_size, uint32_t T,
                           const bpx_pair* scanned_pairs, uint32_t num_pairs,
                           const uint32_t voxels_per_side[3], const float volume_zero_pos[3],
                           const float voxel_inc[9], float t0, float deltaT, int complex_samples,
                           const bpx_options* options, bpx_plan* plan_out);
/* Plan of the GPU path built by the reference pipeline (bp::backproject up to the
 * call_cuda_backprojection call, backproject.hpp:779-891): time limits, transient chunk,
 * pair table, basis and chunked t0 from nlosbpy-style inputs (see bpx_nlos_backproject). */
bpx_status bpx_nlos_plan_create(const float* transient_data, const int64_t* data_shape, int data_ndim,
                                const float* camera_grid_positions, const int64_t* camera_grid_shape,
                                const float* laser_grid_positions, const int64_t* laser_grid_shape,
                                const float camera_position[3], const float laser_position[3],
                                float t0, float deltaT, int is_confocal,
                                const float volume_position[3], const float volume_size[3],
                                const uint32_t voxels_per_side[3], const float wall_normal[3],
                                int assume_row_major, const bpx_options* options,
                                bpx_plan* plan_out);
/* Octree plan (see bpx_octree_backprojection); bpx_plan_execute runs the refinement. */
bpx_status bpx_octree_plan_create(const float* transient_chunk, uint64_t transient_size, uint32_t T,
                                  const bpx_pair* scanned_pairs, uint32_t num_pairs,
                                  const uint32_t voxels_per_side[3], const float volume_zero_pos[3],
                                  const float voxel_inc[9], float t0, float deltaT,
                                  const bpx_octree_options* options, bpx_plan* plan_out);
/* Statistics of the plan's last octree execution. */
bpx_status bpx_octree_plan_stats(bpx_plan plan, bpx_octree_stats* stats);
/* Replaces the thresholds and flags of an octree plan (options->levels must match the plan). */
bpx_status bpx_octree_plan_set_options(bpx_plan plan, const bpx_octree_options* options);
bpx_status bpx_plan_execute(bpx_plan plan, void* hip_stream);
bpx_status bpx_plan_synchronize(bpx_plan plan);
/* Copies the device volume to host (float or interleaved complex). */
bpx_status bpx_plan_read_volume(bpx_plan plan, float* voxel_volume);
/* Describes the plan: kernel path (0 confocal, 1 grid, 2 pairs), voxels per lane (the tile
 * depth, 8 or 16, for the LDS-staged kernel), number of
 * pairs used, chunk length, per-capture scale. */
bpx_status bpx_plan_info(bpx_plan plan, int32_t* path, int32_t* lane_voxels, uint32_t* pairs_used,
                         uint32_t* Tc, float* scale);
/* Name of the gather kernel an execution of the plan launches, with its tile shape (for
 * profiles and reports), as a NUL-terminated string of at most `capacity` bytes.  No reference
 * counterpart (the reference has one kernel per path). */
bpx_status bpx_plan_kernel_name(bpx_plan plan, char* out, uint64_t capacity);
/* The row storage (bpx_storage) the plan's last execution staged its samples in: F16 when an
 * F32 plan's samples were all exact in fp16 (see BPX_FLAG_KEEP_F32_ROWS), else the plan's own.
 * Waits for that execution.  No reference counterpart. */
bpx_status bpx_plan_row_storage(bpx_plan plan, uint32_t* storage);

/* Copies the threshold table of a confocal plan's last execution to host: out[p][j],
 * p < pairs_used, j <= Tc, = the smallest squared wall-point-to-voxel distance whose single-tap
 * bin round((((lw + d) + d) + cw - t0) / deltaT), d = sqrtf(s), is >= j (backproject_cuda.cu:
 * 39-45, 155 for a confocal pair).  The gather kernels pick a voxel's bin with it; exposed so
 * that the table can be checked on its own.  `capacity` = floats available at `out`. */
bpx_status bpx_plan_read_thresholds(bpx_plan plan, float* out, uint64_t capacity);

typedef struct bpx_timing {
    double total_ms;        /* host wall time of all timed executions                    */
    double per_exec_ms;     /* total_ms / iterations                                      */
    double table_kernel_ms; /* mean device time of the staging / threshold-table kernels    */
    double gather_kernel_ms;/* mean device time of the gather kernel per execution         */
    int32_t iterations;
} bpx_timing;
/* Runs `warmup` untimed then `iterations` timed executions on the plan's stream, timing
 * each kernel with HIP events recorded on that stream. */
bpx_status bpx_plan_benchmark(bpx_plan plan, int32_t warmup, int32_t iterations, bpx_timing* out);
/* Autotuner of a confocal plan with fused rows (no reference counterpart; it replaces the
 * occupancy-only launch policy of backproject_cuda.cu:406): executes the plan once plus `reps`
 * times in every gather variant 3-8 that applies to it (the texture-path shapes of 4, 8 and 2
 * voxels per lane; the LDS-staged tile shapes whose bin windows fit) on its own stream and keeps
 * the fastest; every variant gives the same volume bit for bit.  *variant (may be NULL)
 * receives the chosen one; other plans keep theirs. */
bpx_status bpx_plan_tune(bpx_plan plan, int32_t reps, int32_t* variant);
void bpx_plan_destroy(bpx_plan plan);

/* Copies `bytes` (rounded down to 16) between two device buffers with a plain streaming copy
 * kernel `iterations` times; reports the mean kernel time.  The measured HBM bandwidth
 * reference and the calibration of the FETCH_SIZE / WRITE_SIZE counters (no reference
 * counterpart). */
bpx_status bpx_copy_benchmark(uint64_t bytes, int32_t iterations, double* mean_ms);

/* ------------------------------------------------------------------ batched uneven captures
 * Many confocal fp32 captures of different sizes reconstructed by ONE persistent work-queue
 * launch over (capture, tile) items; the per-batch sequence (staging of every capture, queue
 * reset, gather) is a hipGraph replayed by bpx_batch_execute.  Each volume is bit-identical to
 * bpx_backprojection of that capture alone.  No reference counterpart: the reference runs one
 * capture per call_cuda_backprojection (backproject_cuda.cu:370-500); a batch is that call
 * applied to each capture. */
typedef struct bpx_capture_desc {
    const float* transient_chunk;   /* as bpx_backprojection */
    uint64_t transient_size;
    uint32_t T;
    const bpx_pair* scanned_pairs;
    uint32_t num_pairs;
    uint32_t voxels_per_side[3];
    float volume_zero_pos[3];
    float voxel_inc[9];
    float t0, deltaT;
} bpx_capture_desc;
typedef struct bpx_batch_config {
    uint32_t lane_voxels;  /* voxels per lane: 2 or 4                                   */
    uint32_t groups;       /* persistent work-groups (0 = all that are resident at once) */
    uint32_t order;        /* queue: 0 = most work first, 1 = given order, 2 = most work
                              first with texture-path tiles spread among LDS-staged ones */
    uint32_t lds_captures; /* out (get/tune): captures run as LDS-staged tiles (those whose
                              bin windows fit the LDS segment); ignored by set_config   */
    uint32_t lds_depth;    /* LDS-staged tile depth: 8 (16 x 16 x 8) or 16 (16^3); 0 = 8;
                              complex formats always use 8                              */
    uint32_t launches;     /* 1 (or 0): one persistent launch serves the rows' staging, the
                              LDS-staged and the texture-path tiles from one queue; 2: rows
                              staged first, then the LDS-staged queue and a wave-granular
                              texture-path queue as two concurrent persistent launches;
                              3 (real samples): as 2, with the captures that fit 16^3 tiles
                              on a third concurrent queue of 512-lane work-groups (8
                              voxels per lane) and the others' LDS tiles at depth 8    */
} bpx_batch_config;
typedef struct bpx_batch_timing {
    double per_exec_ms;    /* mean device time of one batch execution (HIP events)       */
    double updates_per_s;  /* voxel-pair updates of all captures / per_exec_ms            */
    uint64_t updates;
    int32_t iterations;
} bpx_batch_timing;
typedef struct bpx_batch_s* bpx_batch;
/* options: storage F32, F16 or BF16 for every capture (16-bit rows with a per-capture
 * power-of-two scale, as bpx_backprojection_ex); kahan_block and flags as there, plus
 * BPX_FLAG_COMPLEX_SAMPLES for complex captures (complex64 / complex fp16 / complex bf16 rows). */
bpx_status bpx_batch_create(const bpx_capture_desc* captures, uint32_t count, const bpx_options* options,
                            bpx_batch* out);
uint32_t bpx_batch_size(bpx_batch batch);
/* Autotuner: times every configuration of {2, 4} voxels per lane x {all, half} resident groups
 * x both queue orders on this batch (reps timed executions each) and keeps the fastest. */
bpx_status bpx_batch_tune(bpx_batch batch, int32_t reps, bpx_batch_config* chosen);
bpx_status bpx_batch_set_config(bpx_batch batch, const bpx_batch_config* config);
bpx_status bpx_batch_get_config(bpx_batch batch, bpx_batch_config* config);
bpx_status bpx_batch_execute(bpx_batch batch, void* hip_stream);
bpx_status bpx_batch_synchronize(bpx_batch batch);
bpx_status bpx_batch_read_volume(bpx_batch batch, uint32_t index, float* voxel_volume);
/* Completion time of each capture in the last execution, ms after the batch's queue reset
 * (device wall clock stamped when a capture's last tile finishes). */
bpx_status bpx_batch_latencies(bpx_batch batch, double* ms);
bpx_status bpx_batch_benchmark(bpx_batch batch, int32_t warmup, int32_t iterations, bpx_batch_timing* out);
void bpx_batch_destroy(bpx_batch batch);

/* ------------------------------------------------------------------ learned octree proposals
 * A small 3D conv net (4 layers of 3x3x3 convolutions, 16 channels, BF16 weights and
 * activations on MFMA with FP32 accumulation, ReLU) maps a coarse octree level to one logit
 * per child (8 outputs): "this child contains an occupied voxel".  Inputs per block:
 * U / running max, V / running max, evaluated flag, U / max U, log2(block edge) / 8 (see
 * bpx_octree_options); cells never evaluated are 0.  Attached to
 * an octree plan it narrows the refinement: a child is refined iff its parent's bound passes the
 * fixed threshold AND its probability exceeds proposal_threshold.  No reference counterpart.
 * Parameters: W[4][27][16 out][16 in] then b[4][16], fp32 (bpx_net_param_count() floats). */
typedef struct bpx_net_s* bpx_net;
uint32_t bpx_net_param_count(void);
bpx_status bpx_net_create(uint64_t seed, bpx_net* out);
void bpx_net_destroy(bpx_net net);
bpx_status bpx_net_get_params(bpx_net net, float* params);
bpx_status bpx_net_set_params(bpx_net net, const float* params);
/* input: fp32 [n0][n1][n2][16] (channels >= 5 must be 0); logits out: fp32
 * [n0][n1][n2][16] (channels 0..7 meaningful). */
bpx_status bpx_net_forward(bpx_net net, const float* input, const uint32_t dims[3], float* logits);
/* Weighted BCE loss (sum over voxels with mask != 0, or all when mask is NULL, x 8 outputs)
 * and its gradient w.r.t. every parameter (grads: bpx_net_param_count() floats). */
bpx_status bpx_net_gradients(bpx_net net, const float* input, const uint32_t dims[3],
                             const uint8_t* labels /* [V][8] */, const uint8_t* mask /* [V] */,
                             float pos_weight, double* loss, float* grads);
/* One Adam step on one scene.  octree_plan has been executed (typically with
 * BPX_OCTREE_BOUNDS_ONLY, with or without a net attached); the loss covers the blocks that
 * execution evaluated.  dense_plan has been executed on the same grid; a child's label is
 * (max of its voxels >= occupied_fraction x max of the volume). */
bpx_status bpx_net_train_octree(bpx_net net, bpx_plan octree_plan, bpx_plan dense_plan,
                                float occupied_fraction, float pos_weight, float lr, double* loss);
/* Attach (or, with NULL, detach) a net to an octree plan; the plan does not own it. */
bpx_status bpx_octree_plan_attach_net(bpx_plan plan, bpx_net net);

/* ------------------------------------------------------------------ synthetic captures
 * Renders a transient capture of hidden Lambertian geometry (no reference counterpart).
 * samples: [n][7] = (x, y, z, nx, ny, nz, area) surface samples; normal (0,0,0) = isotropic
 * point.  Pairs follow bpx_pair; path L = lw + |w_l - q| + |q - w_s| + cw, bin
 * floor((L - t0)/deltaT), weight area * cos_l * cos_s / (|w_l - q|^2 |q - w_s|^2) with the
 * cosines taken at the hidden sample (1 for isotropic points).  The rendered expectation is
 * scaled so its maximum bin equals peak_count; if poisson != 0 each bin is replaced by a
 * Poisson draw from a counter-based generator keyed by (seed, pair, bin).  out: [P][T]. */
bpx_status bpx_render_transient(const bpx_pair* pairs, uint32_t num_pairs, uint32_t T, float t0,
                                float deltaT, const float* samples, uint64_t num_samples,
                                float peak_count, int poisson, uint64_t seed, float* out);

#ifdef __cplusplus
}
#endif

#endif /* BPX_H */
