// Host-side internals of libbpx: error reporting, HIP checks, the plan object, and the
// host pipeline restating the reference's bp::backproject / bp::phasor_reconstruction.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "../../include/bpx.h"
#include "bpx_internal.hpp"

namespace bpx {

// ------------------------------------------------------------------ errors
bpx_status set_error(bpx_status s, const std::string& msg);
void clear_error();
struct Failure {
    bpx_status status;
    std::string msg;
};
[[noreturn]] void fail(bpx_status s, const std::string& msg);
void hip_check(hipError_t e, const char* what);
bool device_ready();  // a gfx950 device is visible and selected
// Bound on |approximate - exact quotient| of the table bin lookup for these pairs and grid; the
// lookup is exact below 0.25 (allocates nothing).
double single_tap_error_bound(const bpx_pair* pairs, uint32_t P, const uint32_t n[3], const float zero[3],
                              const float inc[9], float t0, float dT, uint32_t T);
// Per-pair element windows [lo, hi) of a fused confocal plan's rows (2 values per pair), as
// the plan computes them for its staging.
std::vector<uint32_t> confocal_row_windows(const bpx_pair* pairs, uint32_t P, const uint32_t n[3],
                                           const float zero[3], const float inc[9], float t0, float dT,
                                           uint32_t Tc);

// ------------------------------------------------------------------ plan
enum Path : int { kPathConfocal = 0, kPathGrid = 1, kPathPairs = 2, kPathWindow = 3, kPathOctree = 4 };

struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr; o.bytes = 0; }
    void alloc(size_t n);
    void release();
    ~DeviceBuffer() { release(); }
};

struct PlanImpl;

// Octree refinement state (bpx_octree.hip kernels, bpx_octree_plan.cpp orchestration).
struct NetImpl;

struct OctreeState {
    struct Level {
        uint32_t B = 1, n[3] = {0, 0, 0}, W = 0;
        double R = 0;
        float cR = 0;
        DeviceBuffer M;       // window-max table [P][ldm]
        DeviceBuffer list;    // evaluated blocks (packed coordinates)
        DeviceBuffer U, V;    // bound and representative exact value per evaluated block
        DeviceBuffer stats;   // the level's 5 FP64 statistics (see oct_finish_kernel)
        uint32_t count = 0;   // blocks evaluated in the last execution
    };
    uint32_t L = 4, ldm = 0;
    float tau = 0.25f;
    float prob[8] = {0.5f, 0.5f, 0.5f, 0.5f, 0.5f, 0.5f, 0.5f, 0.5f};  // per-level proposal thresholds
    float keep = 0.0f;    // U >= keep x running max overrides the net's veto (0: never)
    float fallback[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // per-level full-resolution fallback probabilities
    uint32_t stride = 1;  // coarse-level pair stride along each wall-grid axis (with a net)
    bool bounds_only = false;
    std::vector<Level> levels;
    DeviceBuffer ping, pong, fine_list, flags, offs, sums, total, stats, partial;
    DeviceBuffer fb_list;                 // low-confidence children of one level (fallback)
    DeviceBuffer fb_voxels;               // their fine voxels, gathered over the levels
    // pair subset of the coarse levels: (x, y, z, c1) and the fused row of each subset pair
    DeviceBuffer sub_pairs, sub_rows;
    uint32_t P_sub = 0, sub_stride = 0;   // subset size and the stride it was built for (0: none)
    uint32_t coarse_P = 0;                // pairs the coarse levels used in the last execution
    std::vector<uint32_t> fallback_voxels;  // fine voxels added by the fallback, per level
    DeviceBuffer bound_part_u, bound_part_v;  // pair-segment partial bounds (kBoundPartCapacity floats)
    static constexpr uint64_t kBoundPartCapacity = 1ull << 20;
    uint32_t* host_total = nullptr;
    double* host_stats = nullptr;
    std::vector<uint32_t> level_active;   // blocks evaluated per level (+ fine voxels last)
    std::vector<double> level_stats;      // 5 FP64 values per level (see oct_finish_kernel)
    NetImpl* net = nullptr;               // learned proposals (not owned), or nullptr
    // second stream (default priority) for the fallback voxels' gather, which runs beside the
    // finer levels' bounds; joined into the plan's stream before the execution ends
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    ~OctreeState();
};

struct PlanImpl {
    hipStream_t stream = nullptr;
    int path = kPathPairs;
    int format = kF32;
    bool cplx = false;
    int variant = 0, K = 4;
    uint32_t P = 0, T = 0, ldh = 0, lds = 0, L = 0, C = 0;
    float t0 = 0, dT = 1, scale = 1;
    Grid g{};
    Tiles tiles{};
    bool sep = false, flat = false, fused = false;
    bool zflat = false;                  // every wall point has the same z (lz), separable grid or not
    bool lds_tiles = false;              // the LDS-staged confocal kernel runs (16 voxels per lane)
    // fp32 rows staged as exact fp16 (FusedH16) whenever every sample of the execution is an
    // integer of magnitude <= 2048; fmt_flag (one word) = 0 then, 1 otherwise
    bool rows16 = false;
    bool keep_f32 = false;               // BPX_FLAG_KEEP_F32_ROWS: rows16 never applies
    DeviceBuffer fmt_flag;
    float lz = 0;
    uint32_t kahan_block = 32;
    float inv2 = 0, inv = 0, c1 = 0;
    // window path
    float vpos[3]{}, vsize_in[3]{}, wzero[3]{}, wvs[3]{}, half_width = 0, c1_lo = 0, c1_hi = 0;
    bool all_confocal = false;
    DeviceBuffer raw, H, S, Z, Z2, pair8, pair4, lasers, spads, vol;
    DeviceBuffer stage;                  // one StageJob (fused confocal layouts)
    // per-pair element windows [lo, hi) of the fused rows that the gathers can read (fused
    // confocal plans, octree plans included; empty = every element staged), device and host copy
    DeviceBuffer win;
    std::vector<uint32_t> win_host;
    std::vector<uint32_t> used;          // source rows of the staged pairs (dropped pairs skipped)
    uint32_t rows_in = 0;                // rows of the capture the plan was made from
    uint32_t stage_zb = 0, stage_sb = 0; // its blocks in the two staging passes
    hipEvent_t ev[8] = {};
    hipEvent_t done = nullptr;           // recorded after each bpx_plan_execute, on its stream
    hipEvent_t ready = nullptr;          // work queued on the plan's own stream (caller-stream runs)
    std::unique_ptr<OctreeState> oct;
    ~PlanImpl();
    void execute(hipStream_t s);
};

// Waits for the plan's own stream and for its last execution on a caller's stream.
void plan_wait(const PlanImpl& pl);
// The storage the plan's last execution staged its rows in (waits for that execution).
int staged_format(const PlanImpl& pl);

void execute_octree(PlanImpl& pl, hipStream_t s);

// The staging job of a fused confocal plan (block offsets 0); launch_stage with the plan's
// device copy stages the plan's rows.
StageJob stage_job(const PlanImpl& pl);
void stage_rows(PlanImpl& pl, hipStream_t s);
// Sets the gather variant of a fused confocal plan (3-8: the same row layout) and what follows
// from it: voxels per lane, the LDS-staged kernel where it applies, exact fp16 rows for its
// shapes of 8 voxels per lane (fp32 storage), and the staging descriptor.
void set_fused_variant(PlanImpl& pl, int variant);

// Proposal net (bpx_net.hip kernels, bpx_net.cpp orchestration).  Four 3x3x3 conv layers of
// 16 channels (inputs zero-padded to 16, 8 output logits = children), ReLU between layers.
struct NetImpl {
    static constexpr int kLayers = 4;
    static constexpr uint32_t kW = 27u * 16u * 16u;          // weights per layer (padded)
    static constexpr uint32_t kParams = kLayers * (kW + 16u); // W[4][27][16][16] then b[4][16]
    static constexpr uint32_t kIn = 5, kOut = 8;              // real input / output channels
    DeviceBuffer params, grads, adam_m, adam_v;   // fp32 [kParams]
    DeviceBuffer wb, wt;                          // bf16 weights and flipped/transposed copies
    DeviceBuffer act[kLayers + 1];                // bf16 activations [V][16]: input, Y1..Y3, (unused)
    DeviceBuffer logits, dlog, dx, dz, labels, mask, loss, vmax, vmax_part, pooled_a, pooled_b;
    static constexpr uint32_t kMaxParts = 2048;   // work-groups of the FP64 max
    uint64_t capacity = 0;                        // voxels the activation buffers hold
    uint64_t step = 0;
    hipStream_t stream = nullptr;
    ~NetImpl();
    void ensure(uint64_t voxels);
    void refresh_bf16(hipStream_t s);
    // forward over a dense grid whose input act[0] is filled; logits fp32 [V][16]
    void forward(uint32_t n0, uint32_t n1, uint32_t n2, hipStream_t s);
    // backward from dlog (bf16 [V][16]); accumulates into grads
    void backward(uint32_t n0, uint32_t n1, uint32_t n2, hipStream_t s);
    void adam(float lr, hipStream_t s);
};

// Proposal hook of the octree: child flags of level l from the net (U >= threshold && logit >
// logit(prob_lo)); returns false when no net is attached.
bool net_child_flags(PlanImpl& pl, uint32_t l, hipStream_t s);

// Describes a capture given to a plan (host buffers).
struct CaptureIn {
    const float* chunk;      // [P][T] real, or interleaved complex
    uint64_t size;           // elements (complex elements when cplx)
    uint32_t T;
    const bpx_pair* pairs;
    uint32_t P;
    float t0, dT;
    bool cplx;
    // Device-side cropping (real samples only): when chunk is NULL and src is set, row p of the
    // chunk is bins [from, from + ncopy) of source row p (pitch src_T) placed at
    // [cfrom, cfrom + ncopy), zero elsewhere -- get_transient_chunk (backproject.hpp:689-740)
    // done by one copy of the source rows and a crop kernel.
    const float* src = nullptr;
    uint32_t src_T = 0, from = 0, cfrom = 0, ncopy = 0;
    // The chunk already in device memory ([P][T] in the layout of `chunk`); the plan takes the
    // buffer over (leaving it empty) when every pair is used, else gathers the used rows.
    DeviceBuffer* device_rows = nullptr;
};

// Replaces a plan's rows with new device rows of the same capture layout ([rows_in][T], the
// layout of CaptureIn::chunk): the plan takes the buffer over when every row is used, else
// gathers the used rows.  The next execution stages and reconstructs them.
void plan_replace_rows(PlanImpl& pl, DeviceBuffer& rows);

// Host -> device copy of a (possibly pageable) host array, enqueued on s and complete when
// this returns; large arrays are pipelined through pinned staging buffers.
void upload_host(void* dst, const void* src, size_t bytes, hipStream_t s);

// Executes a plan, copies its volume to host memory and destroys the plan.
void run_plan_to_host(PlanImpl* plan, float* out);

PlanImpl* make_plan_single_tap(const CaptureIn& c, const uint32_t n[3], const float zero[3],
                               const float inc[9], const bpx_options* opt);
PlanImpl* make_plan_window(const CaptureIn& c, const uint32_t n[3], const float vpos[3],
                           const float vsize[3]);
PlanImpl* make_plan_octree(const CaptureIn& c, const uint32_t n[3], const float zero[3],
                           const float inc[9], const bpx_octree_options* o);
// Validates options against the plan's level count and installs thresholds and flags.
void octree_set_options(OctreeState& st, const bpx_octree_options& o);

}  // namespace bpx

struct bpx_plan_s {
    std::unique_ptr<bpx::PlanImpl> impl;
};

namespace bpx {

// ------------------------------------------------------------------ host pipeline
namespace pipeline {

struct Arrays {
    // row-major copies owned here when a transposition was needed
    std::vector<float> data, cgrid, lgrid;
    const float* data_p = nullptr;
    const float* cgrid_p = nullptr;
    const float* lgrid_p = nullptr;
    std::vector<int64_t> data_shape, cgrid_shape, lgrid_shape;
};

// Normalises inputs to row-major (transposing column-major arrays) and validates shapes.
Arrays normalise(const float* data, const int64_t* data_shape, int data_ndim, const float* cgrid,
                 const int64_t* cgrid_shape, const float* lgrid, const int64_t* lgrid_shape,
                 bool row_major, bool confocal);

// backproject.hpp:356-418
std::vector<bpx_pair> precompute_distances(const float laser_position[3], const float* lgrid,
                                           uint32_t Lx, uint32_t Ly,
                                           const float camera_position[3], const float* cgrid,
                                           uint32_t Cx, uint32_t Cy, bool confocal);
// backproject.hpp:617-674 (row-major); returns (min_T_index - 100, max_T_index + 100)
void time_limits(const float laser_position[3], const float* lgrid, uint32_t Lx, uint32_t Ly,
                 const float camera_position[3], const float* cgrid, uint32_t Cx, uint32_t Cy,
                 const float volume_size[3], const float volume_position[3], float t0, float dT,
                 float* min_idx, float* max_idx);
// backproject.hpp:820-839
// centred: the wall-aligned basis of BPX_FLAG_CENTRED_WALL_BASIS instead of the reference's.
void basis(const float volume_position[3], const float volume_size[3], const uint32_t n[3],
           const float wall_normal[3], float zero[3], float inc[9], bool centred = false);
// backproject.hpp:689-740 (row-major): the chunk's length Tc and the window of every row that
// is copied (source bins [from, from + ncopy) to chunk bins [cfrom, cfrom + ncopy)).
struct ChunkWindow {
    uint32_t Tc, from, cfrom, ncopy;
};
ChunkWindow chunk_window(uint32_t T, float min_idx, float max_idx);
// The same chunk materialised on the host, [rows][Tc].
std::vector<float> transient_chunk(const float* data, uint64_t rows, uint32_t T, float min_idx,
                                   float max_idx, uint32_t* Tc);
// backproject.hpp:430-451 taps (reference parametrisation when sigma <= 0)
int phasor_taps(float wavelength, float dT, uint32_t times, float sigma, std::vector<float>& c,
                std::vector<float>& s);

}  // namespace pipeline

// Phasor FIR of host rows on the device: out [rows][T] interleaved complex (host).
void phasor_filter_host(const float* in, uint64_t rows, uint32_t T, const std::vector<float>& c,
                        const std::vector<float>& s, float* out);

// Synthetic renderer (host, multithreaded).
void render_transient(const bpx_pair* pairs, uint32_t P, uint32_t T, float t0, float dT,
                      const float* samples, uint64_t n, float peak, bool poisson, uint64_t seed,
                      float* out);

}  // namespace bpx
