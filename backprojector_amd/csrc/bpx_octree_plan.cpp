// Host side of the octree plan: level geometry, window-max tables, per-level bound ->
// FP64 reduction -> child selection (fixed threshold or learned proposals), then the exact
// finest level.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>

#include "bpx_host.hpp"

namespace bpx {

static uint32_t round_up32(uint64_t x, uint32_t m) { return (uint32_t)((x + m - 1) / m * m); }

OctreeState::~OctreeState() {
    if (side) (void)hipStreamSynchronize(side);
    if (fork) (void)hipEventDestroy(fork);
    if (join) (void)hipEventDestroy(join);
    if (side) (void)hipStreamDestroy(side);
    if (host_total) (void)hipHostFree(host_total);
    if (host_stats) (void)hipHostFree(host_stats);
}

// Half-diagonal of a block of B fine voxels: max over corner sign patterns of
// |0.5 (B-1) (+-r0 +- r1 +- r2)| with r_k the rows of the grid increment matrix.
static double block_radius(const Grid& g, uint32_t B) {
    double best = 0.0;
    for (int m = 0; m < 4; ++m) {
        const double s1 = (m & 1) ? -1.0 : 1.0, s2 = (m & 2) ? -1.0 : 1.0;
        double v[3];
        for (int k = 0; k < 3; ++k) v[k] = (double)g.inc[k] + s1 * g.inc[3 + k] + s2 * g.inc[6 + k];
        best = std::max(best, std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
    }
    return 0.5 * (double)(B - 1u) * best;
}

void octree_set_options(OctreeState& st, const bpx_octree_options& o) {
    if (o.levels != st.L) fail(BPX_ERR_INVALID_ARGUMENT, "octree levels do not match the plan");
    if (!(o.threshold >= 0.0f) || !std::isfinite(o.threshold))
        fail(BPX_ERR_INVALID_ARGUMENT, "octree threshold must be finite and >= 0");
    auto valid_p = [](float p) { return p > 0.0f && p < 1.0f; };
    if (!valid_p(o.proposal_threshold)) fail(BPX_ERR_INVALID_ARGUMENT, "proposal threshold must lie in (0, 1)");
    float prob[8];
    for (uint32_t l = 0; l < 8; ++l) {
        const float p = o.level_proposal_threshold[l];
        if (p != 0.0f && !valid_p(p))
            fail(BPX_ERR_INVALID_ARGUMENT, "level proposal thresholds must be 0 or lie in (0, 1)");
        prob[l] = p != 0.0f ? p : o.proposal_threshold;
    }
    if (o.flags & ~BPX_OCTREE_BOUNDS_ONLY) fail(BPX_ERR_INVALID_ARGUMENT, "unknown octree flags");
    if (!(o.keep_bound_ratio >= 0.0f) || !std::isfinite(o.keep_bound_ratio))
        fail(BPX_ERR_INVALID_ARGUMENT, "keep_bound_ratio must be finite and >= 0");
    for (uint32_t l = 0; l < 8; ++l) {
        const float f = o.level_fallback_threshold[l];
        if (f != 0.0f && !valid_p(f)) fail(BPX_ERR_INVALID_ARGUMENT, "fallback thresholds must be 0 or lie in (0, 1)");
    }
    if (o.coarse_pair_stride > 64) fail(BPX_ERR_INVALID_ARGUMENT, "coarse_pair_stride must be 0..64");
    st.tau = o.threshold;
    st.keep = o.keep_bound_ratio;
    std::memcpy(st.prob, prob, sizeof(prob));
    std::memcpy(st.fallback, o.level_fallback_threshold, sizeof(st.fallback));
    st.stride = std::max<uint32_t>(1u, o.coarse_pair_stride);
    st.bounds_only = (o.flags & BPX_OCTREE_BOUNDS_ONLY) != 0;
}

// The coarse levels' pair subset for stride s: for a square wall grid (P = side^2 pairs in
// row-major grid order, as the pipeline builds confocal captures) the pairs at every s-th grid
// row and column; otherwise every (s^2)-th pair.  Rebuilt when the stride changes.
static void build_pair_subset(PlanImpl& pl, uint32_t s) {
    OctreeState& st = *pl.oct;
    if (st.sub_stride == s) return;
    const uint32_t P = pl.P;
    std::vector<float> p4((size_t)P * 4);
    if (P) hip_check(hipMemcpy(p4.data(), pl.pair4.ptr, sizeof(float) * p4.size(), hipMemcpyDeviceToHost),
                     "hipMemcpy(pair table)");
    std::vector<uint32_t> rows;
    uint32_t side = (uint32_t)std::lround(std::sqrt((double)P));
    if ((uint64_t)side * side == P && side > 0) {
        for (uint32_t i = 0; i < side; i += s)
            for (uint32_t j = 0; j < side; j += s) rows.push_back(i * side + j);
    } else {
        for (uint64_t p = 0; p < P; p += (uint64_t)s * s) rows.push_back((uint32_t)p);
    }
    std::vector<float> sub(rows.size() * 4);
    for (size_t i = 0; i < rows.size(); ++i) std::memcpy(&sub[4 * i], &p4[4 * (size_t)rows[i]], 16);
    st.sub_pairs.release();
    st.sub_rows.release();
    if (!rows.empty()) {
        st.sub_pairs.alloc(sizeof(float) * sub.size());
        st.sub_rows.alloc(sizeof(uint32_t) * rows.size());
        hip_check(hipMemcpy(st.sub_pairs.ptr, sub.data(), sizeof(float) * sub.size(), hipMemcpyHostToDevice),
                  "hipMemcpy(pair subset)");
        hip_check(hipMemcpy(st.sub_rows.ptr, rows.data(), sizeof(uint32_t) * rows.size(), hipMemcpyHostToDevice),
                  "hipMemcpy(pair subset rows)");
    }
    st.P_sub = (uint32_t)rows.size();
    st.sub_stride = s;
}

PlanImpl* make_plan_octree(const CaptureIn& c, const uint32_t n[3], const float zero[3],
                           const float inc[9], const bpx_octree_options* oo) {
    bpx_octree_options o{4u, 0.25f, 0.5f, 0u, {}, 0.0f, {}, 1u};
    if (oo) o = *oo;
    if (o.levels == 0 || o.levels > 8) fail(BPX_ERR_INVALID_ARGUMENT, "octree levels must be 1..8");
    for (int k = 0; k < 3; ++k) {
        if (n[k] == 0 || n[k] % (1u << o.levels) != 0)
            fail(BPX_ERR_INVALID_ARGUMENT, "voxels_per_side must be divisible by 2^levels");
        if (n[k] > 1024u) fail(BPX_ERR_UNSUPPORTED, "octree volumes up to 1024 voxels per side");
    }
    if (c.cplx) fail(BPX_ERR_UNSUPPORTED, "the octree path takes real samples");
    for (uint32_t p = 0; p < c.P; ++p)
        if (c.pairs && std::memcmp(c.pairs[p].laser_point, c.pairs[p].camera_point, 12) != 0)
            fail(BPX_ERR_UNSUPPORTED, "the octree path takes confocal captures");
    bpx_options po{BPX_STORAGE_F32, 32u, 0u, 5};  // fused layout, 2 voxels per lane
    // the fused rows keep the plain plan's bin windows: the octree reads them only at the bins
    // of fine-grid voxels (fine gathers, fallback voxels, each bound block's representative
    // voxel); the window-max tables of the coarse levels come from the raw rows
    std::unique_ptr<PlanImpl> pl(make_plan_single_tap(c, n, zero, inc, &po));
    auto st = std::make_unique<OctreeState>();
    st->L = o.levels;
    octree_set_options(*st, o);
    const uint32_t P = pl->P, Tc = pl->T;
    st->ldm = round_up32((uint64_t)Tc + 1, 64);
    const uint64_t mbytes = (uint64_t)std::max<uint32_t>(P, 1) * st->ldm * 4;
    if (mbytes >= (1ull << 32)) fail(BPX_ERR_UNSUPPORTED, "window-max tables above 4 GiB per level");
    for (uint32_t l = 0; l < st->L; ++l) {
        OctreeState::Level lv;
        lv.B = 1u << (st->L - l);
        for (int k = 0; k < 3; ++k) lv.n[k] = n[k] / lv.B;
        lv.R = block_radius(pl->g, lv.B);
        lv.W = (uint32_t)std::ceil(4.0 * lv.R / (double)c.dT) + 8u;
        lv.cR = (float)(-3.0 - 2.0 * lv.R / (double)c.dT);
        const uint64_t cells = (uint64_t)lv.n[0] * lv.n[1] * lv.n[2];
        lv.M.alloc(mbytes);
        lv.list.alloc(4 * cells);
        lv.U.alloc(4 * cells);
        lv.V.alloc(4 * cells);
        lv.stats.alloc(8 * sizeof(double));
        st->levels.push_back(std::move(lv));
    }
    if (!winmax_fused_fits(Tc) || st->L > 8) {  // doubling passes through global memory
        st->ping.alloc(mbytes);
        st->pong.alloc(mbytes);
    }
    const uint64_t nvox = (uint64_t)n[0] * n[1] * n[2];
    st->fine_list.alloc(4 * nvox);
    st->fb_list.alloc(4 * nvox);
    st->fb_voxels.alloc(4 * nvox);
    st->flags.alloc(4 * nvox);
    st->offs.alloc(4 * nvox);
    st->sums.alloc(4 * ((nvox + 1023) / 1024 + 1));
    st->total.alloc(2 * sizeof(uint32_t));   // refined children, fallback children
    st->stats.alloc(8 * sizeof(double));
    st->partial.alloc(4 * sizeof(double) * 1024);
    st->bound_part_u.alloc(4 * OctreeState::kBoundPartCapacity);
    st->bound_part_v.alloc(4 * OctreeState::kBoundPartCapacity);
    hip_check(hipStreamCreateWithFlags(&st->side, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipEventCreateWithFlags(&st->fork, hipEventDisableTiming), "hipEventCreate");
    hip_check(hipEventCreateWithFlags(&st->join, hipEventDisableTiming), "hipEventCreate");
    hip_check(hipHostMalloc(&st->host_total, sizeof(uint32_t) * 2, hipHostMallocDefault), "hipHostMalloc");
    hip_check(hipHostMalloc(&st->host_stats, sizeof(double) * 8, hipHostMallocDefault), "hipHostMalloc");
    pl->oct = std::move(st);
    pl->path = kPathOctree;
    return pl.release();
}

// Builds each level's window-max table from the raw rows by doubling widths (in LDS, one pass,
// when a row fits; else through the ping/pong buffers).
static void build_winmax(PlanImpl& pl, uint32_t P, const uint32_t* rows, hipStream_t s) {
    OctreeState& st = *pl.oct;
    const uint32_t Tc = pl.T;
    if (!st.ping.ptr) {
        WinmaxLevels wl{};
        for (uint32_t l = 0; l < st.L; ++l) {
            wl.out[wl.count] = static_cast<float*>(st.levels[l].M.ptr);
            wl.W[wl.count] = st.levels[l].W;
            ++wl.count;
        }
        hip_check(launch_winmax_fused(static_cast<const float*>(pl.raw.ptr), Tc, P, Tc, st.ldm, wl, rows, s),
                  "window-max kernel (LDS)");
        return;
    }
    uint32_t wmax = 1;
    for (const auto& lv : st.levels) wmax = std::max(wmax, lv.W);
    float* cur = static_cast<float*>(st.ping.ptr);
    float* nxt = static_cast<float*>(st.pong.ptr);
    // width 1: max(0, H[b])
    hip_check(launch_winmax_double(static_cast<const float*>(pl.raw.ptr), Tc, cur, st.ldm, P, Tc, 0u, rows, s),
              "window-max kernel");
    uint32_t w = 1;
    for (;;) {
        for (auto& lv : st.levels)
            if (lv.W >= w && lv.W < 2u * w)
                hip_check(launch_winmax_combine(cur, st.ldm, static_cast<float*>(lv.M.ptr), st.ldm, P, Tc,
                                                lv.W - w, s),
                          "window-max combine kernel");
        if (2u * w > wmax) break;
        hip_check(launch_winmax_double(cur, st.ldm, nxt, st.ldm, P, Tc, w, nullptr, s), "window-max kernel");
        std::swap(cur, nxt);
        w *= 2u;
    }
}

void execute_octree(PlanImpl& pl, hipStream_t s) {
    OctreeState& st = *pl.oct;
    const size_t nvox = (size_t)pl.g.n[0] * pl.g.n[1] * pl.g.n[2];
    hip_check(hipGetLastError(), "HIP error pending before the octree execution");
    st.level_active.assign(st.L + 1, 0);
    st.level_stats.assign(5 * st.L, 0.0);
    st.fallback_voxels.assign(st.L, 0);
    for (auto& lv : st.levels) lv.count = 0;
    hip_check(launch_fill(static_cast<float*>(pl.vol.ptr), nvox, 0.0f, s), "fill kernel");
    if (pl.P == 0) return;
    // the coarse levels' pairs: every pair, or the strided subset with a net attached or for
    // training data (bounds only); the bound rule alone keeps every pair, so it stays sound
    const bool subset = (st.net || st.bounds_only) && st.stride > 1;
    if (subset) build_pair_subset(pl, st.stride);
    const uint32_t Pc = subset ? st.P_sub : pl.P;
    const float4* coarse_pairs = static_cast<const float4*>(subset ? st.sub_pairs.ptr : pl.pair4.ptr);
    const uint32_t* coarse_rows = subset ? static_cast<const uint32_t*>(st.sub_rows.ptr) : nullptr;
    st.coarse_P = Pc;
    hip_check(hipEventRecord(pl.ev[0], s), "hipEventRecord");
    void* E = pl.H.ptr;
    const uint32_t e_bytes = (uint32_t)pl.H.bytes;
    stage_rows(pl, s);
    build_winmax(pl, Pc, coarse_rows, s);
    hip_check(hipEventRecord(pl.ev[1], s), "hipEventRecord");
    hip_check(hipMemsetAsync(st.stats.ptr, 0, 8 * sizeof(double), s), "hipMemsetAsync");
    const OctreeState::Level& l0 = st.levels[0];
    uint32_t n = l0.n[0] * l0.n[1] * l0.n[2];
    if (n == 0) fail(BPX_ERR_INVALID_ARGUMENT, "empty coarsest octree level");
    const std::string what = "block list kernel (" + std::to_string(l0.n[0]) + "x" + std::to_string(l0.n[1]) +
                             "x" + std::to_string(l0.n[2]) + ")";
    hip_check(launch_oct_all_blocks(l0.n[0], l0.n[1], l0.n[2], static_cast<uint32_t*>(l0.list.ptr), s),
              what.c_str());
    // full-resolution fallback voxels gathered over the levels (disjoint from the refined ones);
    // each level's are gathered on the side stream as soon as they are known, beside the finer
    // levels' bounds and the refined voxels' gather (different voxels of the same volume)
    uint32_t* fb_voxels = static_cast<uint32_t*>(st.fb_voxels.ptr);
    uint64_t fb_total = 0;
    bool side_used = false;
    OctFineArgs fine{};
    fine.E = E;
    fine.pairs = static_cast<const float4*>(pl.pair4.ptr);
    fine.out = static_cast<float*>(pl.vol.ptr);
    fine.P = pl.P;
    fine.Tc = pl.T;
    fine.lde = pl.ldh;
    fine.e_bytes = e_bytes;
    fine.kahan_block = pl.kahan_block;
    fine.g = pl.g;
    fine.inv2 = pl.inv2;
    fine.lz = pl.lz;
    fine.flat = pl.zflat ? 1u : 0u;
    for (uint32_t l = 0; l < st.L; ++l) {
        OctreeState::Level& lv = st.levels[l];
        lv.count = n;
        st.level_active[l] = n;
        if (n == 0) break;
        const uint32_t* cur = static_cast<const uint32_t*>(lv.list.ptr);
        OctBoundArgs a{};
        a.M = static_cast<const float*>(lv.M.ptr);
        a.E = E;
        a.pairs = coarse_pairs;
        a.erow = coarse_rows;
        a.blocks = cur;
        a.nblocks = n;
        a.U = static_cast<float*>(lv.U.ptr);
        a.V = static_cast<float*>(lv.V.ptr);
        a.P = Pc;
        a.Tc = pl.T;
        a.ldm = st.ldm;
        a.lde = pl.ldh;
        a.m_bytes = (uint32_t)lv.M.bytes;
        a.e_bytes = e_bytes;
        a.g = pl.g;
        a.B = lv.B;
        a.cR = lv.cR;
        a.inv2 = pl.inv2;
        a.lz = pl.lz;
        a.flat = pl.zflat ? 1u : 0u;
        a.partU = static_cast<float*>(st.bound_part_u.ptr);
        a.partV = static_cast<float*>(st.bound_part_v.ptr);
        hip_check(launch_oct_bound(a, oct_bound_segments(n, Pc, OctreeState::kBoundPartCapacity), s),
                  "octree bound kernel");
        hip_check(launch_oct_reduce(a.U, a.V, n, static_cast<double*>(st.partial.ptr), 1024u,
                                    static_cast<double*>(st.stats.ptr), (double)st.tau, s),
                  "octree FP64 reduction");
        hip_check(hipMemcpyAsync(lv.stats.ptr, st.stats.ptr, 5 * sizeof(double), hipMemcpyDeviceToDevice, s),
                  "hipMemcpyAsync");
        // child flags: fixed rule (bound >= tau * running max), narrowed by learned proposals
        // (1 = refine, 2 = full-resolution fallback, 0 = drop)
        if (!net_child_flags(pl, l, s))
            hip_check(launch_oct_child_flags(cur, a.U, n, static_cast<const double*>(st.stats.ptr), nullptr,
                                             lv.n[1], lv.n[2], 0.0f, 0.0f, 0.0f,
                                             static_cast<uint32_t*>(st.flags.ptr), s),
                      "octree child flags");
        uint32_t* next = l + 1 < st.L ? static_cast<uint32_t*>(st.levels[l + 1].list.ptr)
                                      : static_cast<uint32_t*>(st.fine_list.ptr);
        hip_check(launch_oct_scan_children(static_cast<const uint32_t*>(st.flags.ptr), 8u * n, 1u,
                                           static_cast<uint32_t*>(st.offs.ptr),
                                           static_cast<uint32_t*>(st.sums.ptr),
                                           static_cast<uint32_t*>(st.total.ptr), cur, next, s),
                  "octree child scan");
        hip_check(hipMemcpyAsync(st.host_total, st.total.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost, s),
                  "hipMemcpyAsync");
        const bool fallback = st.net && st.fallback[l] > 0.0f;
        if (fallback) {
            hip_check(launch_oct_scan_children(static_cast<const uint32_t*>(st.flags.ptr), 8u * n, 2u,
                                               static_cast<uint32_t*>(st.offs.ptr),
                                               static_cast<uint32_t*>(st.sums.ptr),
                                               static_cast<uint32_t*>(st.total.ptr) + 1, cur,
                                               static_cast<uint32_t*>(st.fb_list.ptr), s),
                      "octree fallback scan");
            hip_check(hipMemcpyAsync(st.host_total + 1, static_cast<uint32_t*>(st.total.ptr) + 1, sizeof(uint32_t),
                                     hipMemcpyDeviceToHost, s),
                      "hipMemcpyAsync");
        }
        hip_check(hipMemcpyAsync(st.host_stats, st.stats.ptr, 5 * sizeof(double), hipMemcpyDeviceToHost, s),
                  "hipMemcpyAsync");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
        for (int k = 0; k < 5; ++k) st.level_stats[5 * l + k] = st.host_stats[k];
        n = st.host_total[0];
        if (fallback && st.host_total[1]) {
            // every fine voxel of each low-confidence child (edge B / 2 fine voxels)
            const uint32_t Bc = lv.B / 2u;
            const uint64_t add = (uint64_t)st.host_total[1] * Bc * Bc * Bc;
            if (fb_total + add > nvox) fail(BPX_ERR_DEVICE, "octree fallback list overflow");
            hip_check(launch_oct_expand(static_cast<const uint32_t*>(st.fb_list.ptr), st.host_total[1], Bc,
                                        fb_voxels + fb_total, s),
                      "octree fallback expansion");
            if (!st.bounds_only) {
                hip_check(hipEventRecord(st.fork, s), "hipEventRecord");
                hip_check(hipStreamWaitEvent(st.side, st.fork, 0), "hipStreamWaitEvent");
                OctFineArgs f = fine;
                f.voxels = fb_voxels + fb_total;
                f.n = (uint32_t)add;
                hip_check(launch_oct_fine(f, st.side), "octree fine kernel (fallback voxels)");
                side_used = true;
            }
            fb_total += add;
            st.fallback_voxels[l] = (uint32_t)add;
        }
    }
    if (st.bounds_only) fb_total = 0;
    if ((uint64_t)n + fb_total > nvox) fail(BPX_ERR_DEVICE, "octree fine list overflow");
    st.level_active[st.L] = st.bounds_only ? 0u : n + (uint32_t)fb_total;
    hip_check(hipEventRecord(pl.ev[2], s), "hipEventRecord");
    if (!st.bounds_only) {
        OctFineArgs f = fine;
        f.voxels = static_cast<const uint32_t*>(st.fine_list.ptr);
        f.n = n;
        hip_check(launch_oct_fine(f, s), "octree fine kernel");
    }
    if (side_used) {
        hip_check(hipEventRecord(st.join, st.side), "hipEventRecord");
        hip_check(hipStreamWaitEvent(s, st.join, 0), "hipStreamWaitEvent");
    }
    hip_check(hipEventRecord(pl.ev[3], s), "hipEventRecord");
}

}  // namespace bpx
