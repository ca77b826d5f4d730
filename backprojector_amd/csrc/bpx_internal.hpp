// Internal declarations shared by the C-ABI layer (host) and the kernel translation unit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace bpx {

// Sample formats of the device-resident transient rows.
enum SampleFormat : int {
    kF32 = 0, kF16 = 1, kBF16 = 2,     // real
    kC64 = 3, kCF16 = 4, kCBF16 = 5,   // complex (interleaved re, im)
};
__host__ __device__ inline bool format_complex(int f) { return f >= kC64; }
inline int format_bytes(int f) {
    switch (f) {
        case kF32: return 4;
        case kF16: case kBF16: return 2;
        case kC64: return 8;
        default: return 4;
    }
}

// Volume grid of the GPU path: v[k] = ((zero[k] + inc[k]x) + inc[3+k]y) + inc[6+k]z.
struct Grid {
    uint32_t n[3];
    float zero[3];
    float inc[9];
};

// Tile decomposition shared by the gather kernels: a work-group of 256 lanes (4 waves)
// covers 8 x 8 x (4K) voxels; each wave a 4 x 4 x 4 lane cube; lane r-th voxel at
// z = tile_z0 + lane_z + 4r.
struct Tiles {
    uint32_t tx, ty, tz, count;
};

// Confocal single-tap kernel parameters (threshold table in squared-distance space).
struct ConfocalArgs {
    const void* H;         // padded rows [P][ldh]: H[p][0] = H[p][Tc+1] = 0, H[p][1+b] = sample b
    const float* S;        // thresholds [P][lds]: S[p][j] = min s with round-bin(s) >= j, j in [0, Tc]
    const float4* pairs;   // (lx, ly, lz, c1): c1 ~ (lw + cw - t0)/dT + 1
    void* out;             // volume, float or float2 per voxel
    uint32_t P, Tc, ldh, lds;
    uint32_t h_bytes, s_bytes;   // buffer sizes for the resource descriptors
    Grid g;
    Tiles t;
    float inv2;            // ~2/deltaT
    float scale;           // per-capture sample scale (power of two)
    float lz;              // common wall z when every pair shares it
    uint32_t kahan_block;
    // rows whose format is decided on the device at each execution: the kernel runs only when
    // *fmt_flag == fmt_want (nullptr: always)
    const uint32_t* fmt_flag;
    uint32_t fmt_want;
};

// Factorised exhaustive-grid kernel parameters (threshold table in path-length space).
// Batched confocal captures: one persistent launch over a queue of staging, LDS-staged and
// texture-path items (batch_queue_kernel).
constexpr uint32_t kStageSpan = 1024;        // fused-row elements per staging chunk (4 per lane)
constexpr uint32_t kBatchStageBlocks = 32;   // fused rows per staging item
enum : uint32_t { kBatchItemStage = 0, kBatchItemLds = 1, kBatchItemTex = 2 };
struct StageJob;
struct BatchQueueArgs {
    const ConfocalArgs* caps;   // per capture; tiles for the batch's voxels per lane
    const uint32_t* flags;      // per capture: bit 0 SEP, bit 1 FLAT
    const StageJob* jobs;       // per capture: its row staging (rows 0: nothing to stage)
    const uint2* items;         // x = capture | type << 28; y = tile, or first row block (stage)
    uint32_t n_items;
    uint32_t spin_limit;        // polls of a capture's staging counter before giving up
    uint32_t* next;             // queue head (0 at launch; the prologue clears next[0..2])
    uint32_t* staged;           // per capture: staging items not finished
    uint32_t* remaining;        // per capture: gather items not finished
    uint64_t* done;             // per capture: wall clock when its last gather item finished
    uint32_t* error;            // nonzero when a wait gave up (never in a correct execution)
};
struct GridArgs {
    const void* H;          // padded rows [L*C][ldh], pair p = l*C + c
    const float4* lasers;   // (x, y, z, laser_wall) per laser point
    const float4* spads;    // (x, y, z, camera_wall) per SPAD point
    const float* Z;         // thresholds [Tc+1] in path-length space
    void* out;
    uint32_t L, C, Tc, ldh;
    uint32_t h_bytes;
    Grid g;
    Tiles t;
    float inv;              // ~1/deltaT
    float c1;               // ~1 - t0/deltaT
    float scale;
    uint32_t kahan_block;
};

// Arbitrary pair list, single tap (no structure assumed).
struct PairsArgs {
    const void* H;
    const float* pairs;     // 8 floats per pair: camera_point[3], laser_point[3], camera_wall, laser_wall
    const float* Z;
    void* out;
    uint32_t P, Tc, ldh;
    uint32_t h_bytes;
    Grid g;
    Tiles t;
    float inv, c1, scale;
    uint32_t kahan_block;
};

// Window (reference CPU path) kernel parameters.
struct WindowArgs {
    const void* PS;         // exclusive prefix sums in double [P][ldp] (complex: double2)
    const float* pairs;     // 8 floats per pair
    const float* Zlo;       // thresholds [T] for the window start bin
    const float* Zhi;       // thresholds [T] for the window end bin
    void* out;
    uint32_t P, T, ldp;
    uint64_t ps_bytes;
    uint32_t n[3];
    float zero[3], vsize[3];
    float inv, c1_lo, c1_hi;
    uint32_t confocal;      // every pair has laser point == camera point
    Tiles t;
};

// Octree level bound: one lane per block (packed x | y << 10 | z << 20).
struct OctBoundArgs {
    const float* M;         // window-max table of the level [P][ldm]
    const void* E;          // fused fp32 sample+threshold rows [P][lde]
    const float4* pairs;    // (x, y, z, c1) of the level's pairs (all, or a subset)
    const uint32_t* erow;   // fused row of pair p in E (nullptr = row p)
    const uint32_t* blocks;
    uint32_t nblocks;
    float* U;               // upper bound per block
    float* V;               // exact value of the block's representative fine voxel
    uint32_t P, Tc, ldm, lde, m_bytes, e_bytes;
    Grid g;                 // fine grid
    uint32_t B;             // block edge in fine voxels
    float cR;               // window-start offset: -3 - 2R/deltaT
    float inv2;
    uint32_t chunk;         // pairs per segment (blockIdx.y); == P when not split
    float* partU;           // [segments][nblocks] partial sums when split, else nullptr
    float* partV;
    float lz;               // wall z of every pair when flat != 0
    uint32_t flat;
};
struct OctFineArgs {
    const void* E;
    const float4* pairs;
    const uint32_t* voxels;  // packed fine voxel coordinates
    uint32_t n;
    float* out;
    uint32_t P, Tc, lde, e_bytes, kahan_block;
    Grid g;
    float inv2;
    float lz;                // wall z of every pair when flat != 0
    uint32_t flat;
};

// ------------------------------------------------------------------ kernel launchers
// All levels' window-max tables in one pass (one work-group per row, row held in LDS); returns
// hipErrorInvalidValue when a row does not fit (callers then use the doubling kernels).
struct WinmaxLevels {
    float* out[8];
    uint32_t W[8];
    uint32_t count;
};
bool winmax_fused_fits(uint32_t Tc);
// row_index: source row of each table row (nullptr = the same row).
hipError_t launch_winmax_fused(const float* in, uint32_t in_pitch, uint32_t rows, uint32_t Tc, uint32_t out_pitch,
                               const WinmaxLevels& lv, const uint32_t* row_index, hipStream_t s);
hipError_t launch_winmax_double(const float* in, uint32_t in_pitch, float* out, uint32_t out_pitch,
                                uint32_t rows, uint32_t Tc, uint32_t half, const uint32_t* row_index, hipStream_t s);
hipError_t launch_winmax_combine(const float* t2, uint32_t pitch2, float* out, uint32_t out_pitch,
                                 uint32_t rows, uint32_t Tc, uint32_t shift, hipStream_t s);
// Pair segments for a level of nblocks blocks (partials must fit part_capacity floats).
uint32_t oct_bound_segments(uint32_t nblocks, uint32_t P, uint64_t part_capacity);
hipError_t launch_oct_bound(const OctBoundArgs& a, uint32_t segments, hipStream_t s);
hipError_t launch_oct_reduce(const float* U, const float* V, uint32_t n, double* partial,
                             uint32_t max_parts, double* stats, double tau, hipStream_t s);
hipError_t launch_oct_fine(const OctFineArgs& a, hipStream_t s);
hipError_t launch_fill(float* p, uint64_t n, float v, hipStream_t s);
hipError_t launch_oct_all_blocks(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t* out, hipStream_t s);
// Children (8 per block, flags[8 b + c]) whose flag equals want, compacted in block order.
hipError_t launch_oct_scan_children(const uint32_t* flags, uint32_t m, uint32_t want, uint32_t* offs, uint32_t* sums,
                                    uint32_t* total, const uint32_t* blocks, uint32_t* children,
                                    hipStream_t s);
hipError_t launch_oct_expand(const uint32_t* blocks, uint32_t n, uint32_t B, uint32_t* out, hipStream_t s);

// ------------------------------------------------------------------ proposal net kernels
hipError_t launch_conv3_fwd(const uint16_t* X, const uint16_t* W, const float* bias, uint16_t* yb,
                            float* yf, uint32_t n0, uint32_t n1, uint32_t n2, int act,
                            uint32_t cout_valid, hipStream_t s);
hipError_t launch_conv3_dw(const uint16_t* dZ, const uint16_t* X, float* dW, float* db, uint32_t n0,
                           uint32_t n1, uint32_t n2, hipStream_t s);
hipError_t launch_flip_transpose(const uint16_t* W, uint16_t* Wt, hipStream_t s);
hipError_t launch_relu_back(const float* dX, const uint16_t* Y, uint16_t* dZ, uint64_t n, hipStream_t s);
hipError_t launch_bce(const float* logits, const uint8_t* labels, const uint8_t* mask, uint32_t V,
                      uint32_t nout, float pos_weight, float scale, uint16_t* dlogits, double* loss,
                      hipStream_t s);
hipError_t launch_adam(float* p, const float* g, float* m, float* v, uint16_t* pb, uint32_t n, uint32_t nb,
                       float lr, float b1, float b2, float eps, float c1, float c2, hipStream_t s);
hipError_t launch_to_bf16(const float* src, uint16_t* dst, uint64_t n, hipStream_t s);
hipError_t launch_oct_features(const uint32_t* blocks, const float* U, const float* V, uint32_t n,
                               const double* stats, uint32_t n1, uint32_t n2, float level_code, uint16_t* X,
                               uint8_t* mask, hipStream_t s);
hipError_t launch_maxpool2(const float* in, float* out, uint32_t m0, uint32_t m1, uint32_t m2, hipStream_t s);
hipError_t launch_child_labels(const float* pooled, uint32_t n0, uint32_t n1, uint32_t n2, const double* vmax,
                               float frac, uint8_t* labels, hipStream_t s);
hipError_t launch_max_f64(const float* x, uint64_t n, double* partial, uint32_t max_parts, double* out,
                          hipStream_t s);
hipError_t launch_oct_child_flags(const uint32_t* blocks, const float* U, uint32_t n, const double* stats,
                                  const float* logits, uint32_t n1, uint32_t n2, float logit_thr, float fallback_thr,
                                  float keep, uint32_t* flags, hipStream_t s);
// All launchers enqueue on `stream` and return the hipError_t of the launch.
// Confocal threshold table S[p][j] (j = 0..Tc) from the pair-independent path-length table
// Z[j] = smallest total path whose single-tap bin is >= j (launch_ztable kind 0, Tc + 1 entries).
hipError_t launch_stable(const float* pair8, uint32_t P, uint32_t Tc, const float* Z,
                         float* S, uint32_t pitch, uint32_t stride, uint32_t offset,
                         hipStream_t stream);

// One confocal capture's staging into a fused sample+threshold layout: Z[0..Tc] first, then
// every row element of the row's window (scaled samples and the pair's threshold slot)
// written once.  A launch serves any number of jobs; each job owns a contiguous range of
// blocks in each of the two passes (z_first: 256 table entries each; s_first: one row each),
// ascending in job order.
struct StageJob {
    const float* src;     // raw samples [rows][Tc] (interleaved re, im for complex formats)
    void* dst;            // fused rows [rows][lde] elements
    const float* pair8;   // pair table, 8 floats per pair
    float* Z;             // [Tc + 1] path-length thresholds (written by the first pass)
    const uint32_t* fmt_flag;   // non-null: rows written in alt_format while *fmt_flag == 0
    // non-null: [rows][2] element windows [lo, hi) of each row that the gathers can read;
    // only those elements are staged (the others keep the zeros the plan wrote once)
    const uint32_t* win;
    uint32_t rows, Tc, lde;
    int32_t format;
    float inv_scale, t0, dT;
    uint32_t z_first, s_first;
    int32_t alt_format;
};
// raw[r][c] = src[map[r]][from + c - cfrom] for c in [cfrom, cfrom + ncopy), else 0
// (r < rows, c < Tc; src rows have pitch src_T; map NULL = identity).
hipError_t launch_crop_rows(const float* src, uint32_t src_T, const uint32_t* map, uint32_t rows, uint32_t Tc,
                            uint32_t from, uint32_t cfrom, uint32_t ncopy, float* raw, hipStream_t stream);
uint32_t stage_z_blocks(const StageJob& j);
uint32_t stage_row_blocks(const StageJob& j);
hipError_t launch_stage(const StageJob* jobs, uint32_t n_jobs, uint32_t z_blocks, uint32_t row_blocks,
                        hipStream_t stream);
bool variant_fused(int variant);
int fused_elem_bytes(int format);
int fused_s_offset(int format);
hipError_t launch_ztable(int kind, float t0, float dT, float half_width, uint32_t n, float* Z,
                         hipStream_t stream);
hipError_t launch_pad_rows(const float* src, uint32_t rows, uint32_t Tc, uint32_t src_pitch,
                           void* dst, uint32_t ldh, int format, float inv_scale,
                           hipStream_t stream);
hipError_t launch_prefix(const float* src, uint32_t rows, uint32_t T, int complex_samples,
                         void* dst, uint32_t ldp, hipStream_t stream);
hipError_t launch_phasor_fir(const float* src, uint32_t rows, uint32_t Tc, const float* taps_c,
                             const float* taps_s, int pulse, void* dst, uint32_t ldh, uint32_t lead,
                             int format, float inv_scale, hipStream_t stream);
hipError_t launch_confocal(const ConfocalArgs& a, int format, int variant, bool sep, bool flat,
                           hipStream_t stream);
hipError_t launch_grid(const GridArgs& a, int format, int variant, hipStream_t stream);
hipError_t launch_pairs(const PairsArgs& a, int format, int variant, hipStream_t stream);
// fp32 samples read straight from unpadded rows [P][Tc] (ldh = Tc); no padded copy needed.
hipError_t launch_pairs_raw(const PairsArgs& a, hipStream_t stream);
// The batch queue kernel (K voxels per lane on the texture path, LDS tile depth kz, storage
// format): resident work-groups, launch, the prologue (path-length tables, queue reset; init =
// [staging items, gather items] per capture), and whether / in how many tiles a capture runs
// LDS-staged.
// launches 1: the queue serves every kind; 2: LDS-staged tiles only, beside the wave-granular
// texture-path queue (batch_tex_kernel, items (capture, tile * 4 + quarter), head next[1]).
int batch_queue_resident_groups(int K, int kz, int format, int launches);
int batch_lds_depth(int kz, int format);   // the depth the queue kernel uses for this format
hipError_t launch_batch_queue(const BatchQueueArgs& b, int K, int kz, int format, int launches, uint32_t groups,
                              hipStream_t s);
int batch_tex_resident_groups(int K, int format);
// launches 3: the 16^3-tile LDS-staged queue of 512-lane work-groups (real formats; head next[2])
int batch_wide_resident_groups(int format);
hipError_t launch_batch_wide(const BatchQueueArgs& b, int format, uint32_t groups, hipStream_t s);
hipError_t launch_batch_tex(const BatchQueueArgs& b, int K, int format, uint32_t groups, hipStream_t s);
hipError_t launch_batch_prologue(const StageJob* jobs, uint32_t n_jobs, uint32_t z_blocks, const BatchQueueArgs& b,
                                 const uint32_t* init, uint32_t n, uint64_t* start, hipStream_t s);
bool batch_lds_applies(const Grid& g, float inv2, int format, bool sep, bool flat, int kz);
uint32_t batch_lds_tiles(const Grid& g, int kz);
hipError_t launch_window(const WindowArgs& a, int complex_samples, hipStream_t stream);
hipError_t launch_abs_max(const float* src, uint64_t n, float* result_dev, hipStream_t stream);
// *flag |= 1 unless every one of n samples is an integer of magnitude <= 2048 (exact in fp16);
// the caller clears *flag first.
hipError_t launch_exact16_check(const float* src, uint64_t n, uint32_t* flag, hipStream_t stream);
// |v| and atan2(im, re) of n interleaved complex values (either output may be null).
hipError_t launch_complex_post(const void* complex_volume, uint64_t n, float* mag, float* phase,
                               hipStream_t stream);
hipError_t launch_copy(void* dst, const void* src, uint64_t bytes, hipStream_t stream);

// Voxels per lane (K) of each tuned variant; variant -1 picks the default.
constexpr int kDefaultVariant = 5;  // fused layout, 2 voxels per lane (where the LDS kernel does not apply)
constexpr int kLdsVariant = 6;      // LDS-staged 16^3 tiles, 256 lanes (fused rows, separable, flat);
                                    // + 1: 16 x 16 x 8 tiles, 256 lanes; + 2: 16^3 tiles, 512 lanes
constexpr uint64_t kLdsDeepTiles = 1024;  // 16^3 tiles of 256 lanes by default from this many up
constexpr uint64_t kLdsMinTiles = 512;    // the other LDS shapes by default from this many up
// Whether the LDS-staged kernel's per-pair bin windows over 16 x 16 x kz tiles fit a segment of
// seg elements for this grid spacing; the segment length of a storage format; the tile depth a
// variant (6, 7 or 8) gives for a storage format (complex 16-bit formats: 16 x 16 x 8 tiles;
// complex64 runs as two real passes with the real formats' shapes).
bool confocal_lds_fits(const float inc[9], float inv2, int kz, uint32_t seg);
uint32_t lds_seg_elems(int format);
// Tile of an LDS-staged variant (6, 7, 8) for a storage format: kz voxels per lane along z,
// nw waves per work-group (= pairs per stage); tiles are 16 x 16 x depth().
struct LdsShape {
    int kz, nw;
    int depth() const { return kz * nw / 4; }
};
LdsShape lds_shape(int variant, int format);
int variant_k(int variant);
int variant_count();

// Tiles of a volume for a given K.
Tiles make_tiles(const uint32_t n[3], int K);

}  // namespace bpx
