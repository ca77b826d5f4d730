// Batched uneven confocal captures: one device-resident plan per capture and one persistent
// launch over a single work queue of staging items (fused rows and thresholds), LDS-staged
// tiles (captures whose bin windows fit the LDS segment) and texture-path tiles (the rest),
// after a prologue launch that builds every capture's path-length table and resets the queue.
// The queue holds every staging item first and then the gather items; a capture's gathers wait
// for its staging, so the first captures' gathers overlap the last captures' staging.  The sequence is captured in a hipGraph that bpx_batch_execute replays.
// A project-written autotuner times launch configurations (texture-path voxels per lane, LDS
// tile depth, queue order, resident work-groups) on the batch itself and keeps the fastest.
// No reference counterpart: the reference reconstructs one capture per call
// (call_cuda_backprojection, backproject_cuda.cu:370-500), with a launch shape chosen from
// occupancy alone (cudaOccupancyMaxPotentialBlockSize, :406).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "bpx_host.hpp"

namespace bpx {

struct BatchImpl {
    std::vector<std::unique_ptr<PlanImpl>> plans;
    std::vector<uint64_t> updates;   // voxel-pair updates per capture
    hipStream_t stream = nullptr;
    hipEvent_t ev[2] = {};
    hipEvent_t done = nullptr, ready = nullptr;   // as PlanImpl::done / ready
    DeviceBuffer caps, flags, items, items_tex, items_wide, jobs, init, next, staged, remaining, error;
    DeviceBuffer stamps;   // uint64 [1 + captures]: batch start, then each capture's completion
    DeviceBuffer zjobs;    // StageJob per capture with pairs (path-length tables, prologue)
    uint32_t n_zjobs = 0, z_blocks = 0, row_blocks = 0;
    uint32_t n_items = 0, n_tex_items = 0, n_wide_items = 0;
    uint32_t tex_groups = 0, wide_groups = 0;
    // two- and three-launch forms: the texture-path queue (and the 16^3-tile queue) run beside
    // the LDS-staged one on other streams (default priority), forked from and joined back into
    // the captured sequence
    hipStream_t side = nullptr, side2 = nullptr;
    hipEvent_t fork = nullptr, join = nullptr, join2 = nullptr;
    int format = kF32;     // storage of every capture (fused rows)
    bool cplx = false;     // complex captures and volumes
    bpx_batch_config cfg{2u, 0u, 0u, 0u, 8u, 1u};
    uint32_t groups = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double wall_rate_khz = 0.0;
    ~BatchImpl() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (done) (void)hipEventDestroy(done);
        if (ready) (void)hipEventDestroy(ready);
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        if (join2) (void)hipEventDestroy(join2);
        if (side) (void)hipStreamDestroy(side);
        if (side2) (void)hipStreamDestroy(side2);
        if (stream) (void)hipStreamDestroy(stream);
    }
    void build(const bpx_batch_config& c);
    // Waits for the batch's own stream and for its last execution on a caller's stream, then
    // checks the queue's error word.
    void wait() const {
        hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        hip_check(hipEventSynchronize(done), "hipEventSynchronize");
        uint32_t e = 0;
        hip_check(hipMemcpy(&e, error.ptr, 4, hipMemcpyDeviceToHost), "hipMemcpy(error word)");
        if (e) fail(BPX_ERR_DEVICE, "batch queue: a capture's staging did not finish in time");
    }
};

void BatchImpl::build(const bpx_batch_config& c) {
    if (c.lane_voxels != 2u && c.lane_voxels != 4u) fail(BPX_ERR_INVALID_ARGUMENT, "lane_voxels must be 2 or 4");
    if (c.order > 2u) fail(BPX_ERR_INVALID_ARGUMENT, "order must be 0, 1 or 2");
    const uint32_t kz_req = c.lds_depth ? c.lds_depth : 8u;
    if (kz_req != 8u && kz_req != 16u) fail(BPX_ERR_INVALID_ARGUMENT, "lds_depth must be 8 or 16");
    const uint32_t launches = c.launches ? c.launches : 1u;
    if (launches > 3u) fail(BPX_ERR_INVALID_ARGUMENT, "launches must be 1, 2 or 3");
    const bool three = launches == 3u;
    if (three && format_complex(format)) fail(BPX_ERR_INVALID_ARGUMENT, "launches = 3 needs real samples");
    const int K = (int)c.lane_voxels;
    const int kz = three ? 8 : batch_lds_depth((int)kz_req, format);
    const bool two = launches >= 2u;   // rows staged first, the texture-path tiles on their own queue
    const uint32_t n = (uint32_t)plans.size();
    // per-capture arguments, tiles and staging jobs for this configuration
    std::vector<ConfocalArgs> args(n);
    std::vector<StageJob> sj(std::max<uint32_t>(n, 1));
    std::vector<uint32_t> fl(n), nt(n), ns(n);
    std::vector<char> lds(n, 0), wide(n, 0);
    for (uint32_t i = 0; i < n; ++i) {
        const PlanImpl& pl = *plans[i];
        ConfocalArgs& a = args[i];
        a.H = pl.H.ptr;
        a.pairs = static_cast<const float4*>(pl.pair4.ptr);
        a.out = pl.vol.ptr;
        a.P = pl.P;
        a.Tc = pl.T;
        a.ldh = pl.ldh;
        a.h_bytes = (uint32_t)pl.H.bytes;
        a.g = pl.g;
        a.t = make_tiles(pl.g.n, K);
        a.inv2 = pl.inv2;
        a.scale = pl.scale;
        a.lz = pl.lz;
        a.kahan_block = pl.kahan_block;
        fl[i] = (pl.sep ? 1u : 0u) | (pl.sep && pl.flat ? 2u : 0u);
        // captures whose bin windows fit the LDS segment run LDS-staged tiles, the others
        // texture-path tiles
        // (three launches: those that fit 16^3 tiles take the 512-lane 16^3-tile queue)
        wide[i] = three && pl.P && batch_lds_applies(pl.g, pl.inv2, format, pl.sep, pl.flat, 16);
        lds[i] = pl.P && (wide[i] || batch_lds_applies(pl.g, pl.inv2, format, pl.sep, pl.flat, kz));
        // (two or three launches: a texture-path tile is 4 wave items, one per quarter)
        nt[i] = !pl.P ? 0u
                : wide[i] ? batch_lds_tiles(pl.g, 16)
                : lds[i]  ? batch_lds_tiles(pl.g, kz)
                          : (two ? 4u : 1u) * a.t.count;
        if (pl.P) {
            sj[i] = stage_job(pl);
            ns[i] = two ? 0u : (stage_row_blocks(sj[i]) + kBatchStageBlocks - 1u) / kBatchStageBlocks;
        }
    }
    // capture order: 0 = most work first, 1 = given order, 2 = most work first with the
    // texture-path tiles spread evenly among the LDS-staged ones (the two kinds load different
    // units -- texture addressing against VALU and LDS -- so work-groups of both kinds on one CU
    // overlap).  Every staging item comes first (capture order), then the gather items.
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    if (c.order != 1)
        std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return updates[x] > updates[y]; });
    std::vector<uint32_t> it;
    auto push_stage = [&](uint32_t i) {
        for (uint32_t k = 0; k < ns[i]; ++k) {
            it.push_back(i | (kBatchItemStage << 28));
            it.push_back(k * kBatchStageBlocks);
        }
    };
    auto push_gather = [&](uint32_t i) {
        const uint32_t type = lds[i] ? kBatchItemLds : kBatchItemTex;
        for (uint32_t t = 0; t < nt[i]; ++t) {
            it.push_back(i | (type << 28));
            it.push_back(t);
        }
    };
    for (uint32_t k = 0; k < n; ++k) push_stage(order[k]);
    if (c.order == 2) {
        std::vector<uint32_t> a, b;   // LDS-staged and texture-path gather items, capture order
        for (uint32_t k = 0; k < n; ++k) {
            const uint32_t i = order[k];
            std::vector<uint32_t>& q = lds[i] ? a : b;
            const uint32_t type = lds[i] ? kBatchItemLds : kBatchItemTex;
            for (uint32_t t = 0; t < nt[i]; ++t) {
                q.push_back(i | (type << 28));
                q.push_back(t);
            }
        }
        // proportional merge: after each item of the longer list, the share of the other
        const size_t na = a.size() / 2, nb = b.size() / 2;
        size_t ia = 0, ib = 0;
        while (ia < na || ib < nb) {
            const bool take_a = ib >= nb || (ia < na && (ia + 1) * nb <= (ib + 1) * na);
            std::vector<uint32_t>& q = take_a ? a : b;
            size_t& k = take_a ? ia : ib;
            it.push_back(q[2 * k]);
            it.push_back(q[2 * k + 1]);
            ++k;
        }
    } else {
        for (uint32_t k = 0; k < n; ++k) push_gather(order[k]);
    }
    // two launches: the texture-path items go to their own (wave-granular) queue; three: the
    // 16^3 tiles to theirs as well
    std::vector<uint32_t> itx, itw;
    if (two) {
        std::vector<uint32_t> keep;
        for (size_t k = 0; k < it.size(); k += 2) {
            const uint32_t ci = it[k] & 0x0fffffffu;
            std::vector<uint32_t>& q = (it[k] >> 28) == kBatchItemTex ? itx : wide[ci] ? itw : keep;
            q.push_back(it[k]);
            q.push_back(it[k + 1]);
        }
        it.swap(keep);
    }
    n_items = (uint32_t)(it.size() / 2);
    n_tex_items = (uint32_t)(itx.size() / 2);
    n_wide_items = (uint32_t)(itw.size() / 2);
    // path-length tables of every capture in the prologue: each job's blocks follow the previous one's
    std::vector<StageJob> zj;
    uint64_t zb = 0, rb = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (!plans[i]->P) continue;
        StageJob j = sj[i];
        j.z_first = (uint32_t)zb;
        j.s_first = (uint32_t)rb;
        zb += stage_z_blocks(j);
        rb += stage_row_blocks(j);
        zj.push_back(j);
    }
    if (zb >= (1ull << 31) || rb >= (1ull << 31)) fail(BPX_ERR_UNSUPPORTED, "batch staging too large");
    n_zjobs = (uint32_t)zj.size();
    z_blocks = (uint32_t)zb;
    row_blocks = (uint32_t)rb;
    std::vector<uint32_t> in(2 * std::max<uint32_t>(n, 1), 0u);
    for (uint32_t i = 0; i < n; ++i) {
        in[2 * i] = ns[i];
        in[2 * i + 1] = nt[i];
    }
    zjobs.alloc(sizeof(StageJob) * std::max<size_t>(zj.size(), 1));
    jobs.alloc(sizeof(StageJob) * sj.size());
    init.alloc(4u * in.size());
    caps.alloc(sizeof(ConfocalArgs) * std::max<uint32_t>(n, 1));
    flags.alloc(4u * std::max<uint32_t>(n, 1));
    items.alloc(8u * std::max<uint32_t>(n_items, 1));
    items_tex.alloc(8u * std::max<uint32_t>(n_tex_items, 1));
    items_wide.alloc(8u * std::max<uint32_t>(n_wide_items, 1));
    if (n_zjobs)
        hip_check(hipMemcpy(zjobs.ptr, zj.data(), sizeof(StageJob) * n_zjobs, hipMemcpyHostToDevice), "hipMemcpy");
    hip_check(hipMemcpy(jobs.ptr, sj.data(), sizeof(StageJob) * sj.size(), hipMemcpyHostToDevice), "hipMemcpy");
    hip_check(hipMemcpy(init.ptr, in.data(), 4u * in.size(), hipMemcpyHostToDevice), "hipMemcpy");
    if (n) {
        hip_check(hipMemcpy(caps.ptr, args.data(), sizeof(ConfocalArgs) * n, hipMemcpyHostToDevice), "hipMemcpy");
        hip_check(hipMemcpy(flags.ptr, fl.data(), 4u * n, hipMemcpyHostToDevice), "hipMemcpy");
    }
    if (n_items) hip_check(hipMemcpy(items.ptr, it.data(), 8u * n_items, hipMemcpyHostToDevice), "hipMemcpy");
    if (n_tex_items)
        hip_check(hipMemcpy(items_tex.ptr, itx.data(), 8u * n_tex_items, hipMemcpyHostToDevice), "hipMemcpy");
    if (n_wide_items)
        hip_check(hipMemcpy(items_wide.ptr, itw.data(), 8u * n_wide_items, hipMemcpyHostToDevice), "hipMemcpy");
    const int resident = batch_queue_resident_groups(K, kz, format, (int)launches);
    if (resident <= 0) fail(BPX_ERR_DEVICE, "cannot query the occupancy of the batch kernel");
    groups = c.groups ? std::min<uint32_t>(c.groups, (uint32_t)resident) : (uint32_t)resident;
    groups = std::max<uint32_t>(1u, std::min<uint32_t>(groups, std::max<uint32_t>(n_items, 1u)));
    wide_groups = 0;
    if (n_wide_items) {
        const int rw = batch_wide_resident_groups(format);
        if (rw <= 0) fail(BPX_ERR_DEVICE, "cannot query the occupancy of the batch 16^3-tile kernel");
        wide_groups = c.groups ? std::min<uint32_t>(c.groups, (uint32_t)rw) : (uint32_t)rw;
        wide_groups = std::max<uint32_t>(1u, std::min<uint32_t>(wide_groups, n_wide_items));
    }
    if (two && n_tex_items) {
        const int rt = batch_tex_resident_groups(K, format);
        if (rt <= 0) fail(BPX_ERR_DEVICE, "cannot query the occupancy of the batch texture-path kernel");
        tex_groups = c.groups ? std::min<uint32_t>(c.groups, (uint32_t)rt) : (uint32_t)rt;
        tex_groups = std::max<uint32_t>(1u, std::min<uint32_t>(tex_groups, (n_tex_items + 3u) / 4u));
    }
    cfg = c;
    cfg.groups = groups;
    cfg.lds_depth = (uint32_t)kz;
    cfg.launches = launches;
    cfg.lds_captures = (uint32_t)std::count(lds.begin(), lds.end(), (char)1);   // 16^3 tiles included
    BatchQueueArgs q{};
    q.caps = static_cast<const ConfocalArgs*>(caps.ptr);
    q.flags = static_cast<const uint32_t*>(flags.ptr);
    q.jobs = static_cast<const StageJob*>(jobs.ptr);
    q.items = static_cast<const uint2*>(items.ptr);
    q.n_items = n_items;
    q.spin_limit = 1u << 24;
    q.next = static_cast<uint32_t*>(next.ptr);
    q.staged = static_cast<uint32_t*>(staged.ptr);
    q.remaining = static_cast<uint32_t*>(remaining.ptr);
    q.done = static_cast<uint64_t*>(stamps.ptr) + 1;
    q.error = static_cast<uint32_t*>(error.ptr);
    // the per-batch sequence as one graph (no memset nodes: the reset is a kernel)
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    BatchQueueArgs qt = q;   // the texture-path queue of the two-launch form
    qt.items = static_cast<const uint2*>(items_tex.ptr);
    qt.n_items = n_tex_items;
    qt.next = static_cast<uint32_t*>(next.ptr) + 1;
    BatchQueueArgs qw = q;   // the 16^3-tile queue of the three-launch form
    qw.items = static_cast<const uint2*>(items_wide.ptr);
    qw.n_items = n_wide_items;
    qw.next = static_cast<uint32_t*>(next.ptr) + 2;
    hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    bool side_joined = false, side2_joined = false;   // side streams in the capture, not yet joined back
    try {
        for (auto& pl : plans) {
            if (pl->P != 0) continue;
            const size_t nvox = (size_t)pl->g.n[0] * pl->g.n[1] * pl->g.n[2];
            hip_check(launch_fill(static_cast<float*>(pl->vol.ptr), nvox * (cplx ? 2u : 1u), 0.0f, stream),
                      "fill kernel");
        }
        hip_check(launch_batch_prologue(static_cast<const StageJob*>(zjobs.ptr), n_zjobs, z_blocks, q,
                                        static_cast<const uint32_t*>(init.ptr), n,
                                        static_cast<uint64_t*>(stamps.ptr), stream),
                  "batch prologue kernel");
        if (two) {
            // every row staged first (no waits in either queue), then both queues side by side
            hip_check(launch_stage(static_cast<const StageJob*>(zjobs.ptr), n_zjobs, 0u, row_blocks, stream),
                      "batch row staging kernel");
            // every queue with items on its own stream: the 16 x 16 x kz one on the graph's
            // stream, the texture-path one on `side`, the 16^3 one on `side2`
            const bool fork_side = n_tex_items && (n_items || n_wide_items);
            const bool fork_side2 = n_wide_items && (n_items || n_tex_items);
            if (fork_side || fork_side2) hip_check(hipEventRecord(fork, stream), "hipEventRecord");
            if (fork_side) {
                hip_check(hipStreamWaitEvent(side, fork, 0), "hipStreamWaitEvent");
                side_joined = true;
            }
            if (fork_side2) {
                hip_check(hipStreamWaitEvent(side2, fork, 0), "hipStreamWaitEvent");
                side2_joined = true;
            }
            if (n_items)
                hip_check(launch_batch_queue(q, K, kz, format, (int)launches, groups, stream),
                          "batch LDS-staged queue");
            if (n_tex_items)
                hip_check(launch_batch_tex(qt, K, format, tex_groups, fork_side ? side : stream),
                          "batch texture-path queue");
            if (n_wide_items)
                hip_check(launch_batch_wide(qw, format, wide_groups, fork_side2 ? side2 : stream),
                          "batch 16^3-tile queue");
            if (fork_side) {
                hip_check(hipEventRecord(join, side), "hipEventRecord");
                hip_check(hipStreamWaitEvent(stream, join, 0), "hipStreamWaitEvent");
                side_joined = false;
            }
            if (fork_side2) {
                hip_check(hipEventRecord(join2, side2), "hipEventRecord");
                hip_check(hipStreamWaitEvent(stream, join2, 0), "hipStreamWaitEvent");
                side2_joined = false;
            }
        } else if (n_items) {
            hip_check(launch_batch_queue(q, K, kz, format, 1, groups, stream), "batch queue kernel");
        }
    } catch (...) {
        // a side stream left in the capture would make the capture fail as unjoined and stay in
        // capture mode: join it back first, then end (and drop) the capture
        if (side_joined && hipEventRecord(join, side) == hipSuccess) (void)hipStreamWaitEvent(stream, join, 0);
        if (side2_joined && hipEventRecord(join2, side2) == hipSuccess) (void)hipStreamWaitEvent(stream, join2, 0);
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    hip_check(hipStreamEndCapture(stream, &graph), "hipStreamEndCapture");
    hip_check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
}

}  // namespace bpx

using namespace bpx;

struct bpx_batch_s {
    std::unique_ptr<BatchImpl> impl;
};

namespace {
template <class F>
bpx_status batch_guard(F&& f) {
    clear_error();
    try {
        f();
        return BPX_OK;
    } catch (const Failure& e) {
        return set_error(e.status, e.msg);
    } catch (const std::bad_alloc&) {
        return set_error(BPX_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
}

double run_timed(BatchImpl& b, int32_t iters) {
    hip_check(hipEventRecord(b.ev[0], b.stream), "hipEventRecord");
    for (int32_t i = 0; i < iters; ++i) hip_check(hipGraphLaunch(b.exec, b.stream), "hipGraphLaunch");
    hip_check(hipEventRecord(b.ev[1], b.stream), "hipEventRecord");
    hip_check(hipEventSynchronize(b.ev[1]), "hipEventSynchronize");
    float ms = 0.0f;
    hip_check(hipEventElapsedTime(&ms, b.ev[0], b.ev[1]), "hipEventElapsedTime");
    return (double)ms / std::max(iters, 1);
}
}  // namespace

extern "C" {

bpx_status bpx_batch_create(const bpx_capture_desc* captures, uint32_t count, const bpx_options* options,
                            bpx_batch* out) {
    return batch_guard([&] {
        if (!out) fail(BPX_ERR_INVALID_ARGUMENT, "NULL output");
        *out = nullptr;
        if (count && !captures) fail(BPX_ERR_INVALID_ARGUMENT, "NULL captures");
        if (!device_ready()) fail(BPX_ERR_NO_DEVICE, "no gfx950 device available");
        bpx_options o{BPX_STORAGE_F32, 32u, 0u, -1};
        if (options) o = *options;
        if (o.storage > BPX_STORAGE_BF16) fail(BPX_ERR_INVALID_ARGUMENT, "unknown storage format");
        o.variant = 5;  // fused rows: the layout the batch kernel reads
        o.flags |= BPX_FLAG_KEEP_F32_ROWS;   // the batch stages the rows itself, in one format
        auto b = std::make_unique<BatchImpl>();
        b->cplx = (o.flags & BPX_FLAG_COMPLEX_SAMPLES) != 0;
        o.flags &= ~BPX_FLAG_COMPLEX_SAMPLES;
        if (b->cplx)
            b->format = o.storage == BPX_STORAGE_F16 ? kCF16 : o.storage == BPX_STORAGE_BF16 ? kCBF16 : kC64;
        else
            b->format = o.storage == BPX_STORAGE_F16 ? kF16 : o.storage == BPX_STORAGE_BF16 ? kBF16 : kF32;
        hip_check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "hipEventCreate");
        hip_check(hipEventCreateWithFlags(&b->join, hipEventDisableTiming), "hipEventCreate");
        hip_check(hipStreamCreateWithFlags(&b->side2, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipEventCreateWithFlags(&b->join2, hipEventDisableTiming), "hipEventCreate");
        for (auto& e : b->ev) hip_check(hipEventCreate(&e), "hipEventCreate");
        hip_check(hipEventCreateWithFlags(&b->done, hipEventDisableTiming), "hipEventCreate");
        hip_check(hipEventCreateWithFlags(&b->ready, hipEventDisableTiming), "hipEventCreate");
        int dev = 0, rate = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        hip_check(hipDeviceGetAttribute(&rate, hipDeviceAttributeWallClockRate, dev), "hipDeviceGetAttribute");
        b->wall_rate_khz = rate > 0 ? (double)rate : 1.0e5;
        for (uint32_t i = 0; i < count; ++i) {
            const bpx_capture_desc& d = captures[i];
            CaptureIn c{d.transient_chunk, d.transient_size, d.T, d.scanned_pairs, d.num_pairs, d.t0, d.deltaT,
                        b->cplx};
            std::unique_ptr<PlanImpl> pl(make_plan_single_tap(c, d.voxels_per_side, d.volume_zero_pos, d.voxel_inc, &o));
            if (pl->P && (pl->path != kPathConfocal || !pl->fused || pl->format != b->format))
                fail(BPX_ERR_UNSUPPORTED, "capture " + std::to_string(i) + " is not a confocal capture");
            b->updates.push_back((uint64_t)pl->P * pl->g.n[0] * pl->g.n[1] * pl->g.n[2]);
            b->plans.push_back(std::move(pl));
        }
        const uint32_t n = std::max<uint32_t>(count, 1);
        b->next.alloc(12);   // queue heads: the (LDS-staged) queue, texture-path, 16^3-tile
        b->error.alloc(4);
        hip_check(hipMemset(b->error.ptr, 0, 4), "hipMemset");
        b->staged.alloc(4u * n);
        b->remaining.alloc(4u * n);
        b->stamps.alloc(8u * (n + 1u));
        b->build(bpx_batch_config{2u, 0u, 0u, 0u, 8u, 1u});
        auto h = std::make_unique<bpx_batch_s>();
        h->impl = std::move(b);
        *out = h.release();
    });
}

void bpx_batch_destroy(bpx_batch batch) {
    if (!batch) return;
    (void)hipStreamSynchronize(batch->impl->stream);
    if (batch->impl->done) (void)hipEventSynchronize(batch->impl->done);
    delete batch;
}

bpx_status bpx_batch_set_config(bpx_batch batch, const bpx_batch_config* config) {
    return batch_guard([&] {
        if (!batch || !config) fail(BPX_ERR_INVALID_ARGUMENT, "NULL argument");
        batch->impl->wait();
        batch->impl->build(*config);
    });
}

bpx_status bpx_batch_get_config(bpx_batch batch, bpx_batch_config* config) {
    return batch_guard([&] {
        if (!batch || !config) fail(BPX_ERR_INVALID_ARGUMENT, "NULL argument");
        *config = batch->impl->cfg;
    });
}

bpx_status bpx_batch_tune(bpx_batch batch, int32_t reps, bpx_batch_config* chosen) {
    return batch_guard([&] {
        if (!batch) fail(BPX_ERR_INVALID_ARGUMENT, "NULL batch");
        BatchImpl& b = *batch->impl;
        b.wait();
        reps = std::max(reps, 1);
        bpx_batch_config best{2u, 0u, 0u, 0u, 8u, 1u};
        double best_ms = INFINITY;
        for (uint32_t launches : {1u, 2u, 3u})
            for (uint32_t k : {2u, 4u})
                for (uint32_t kz : {8u, 16u}) {
                    if (kz == 16u && format_complex(b.format)) continue;   // complex formats: depth 8
                    if (launches == 3u && (kz == 16u || format_complex(b.format))) continue;   // both depths
                    const int resident = batch_queue_resident_groups((int)k, (int)kz, b.format, (int)launches);
                    for (uint32_t frac : {1u, 2u})
                        for (uint32_t order : {0u, 1u, 2u}) {
                            const bpx_batch_config c{k, resident > 0 ? std::max(1u, (uint32_t)resident / frac) : 0u,
                                                     order, 0u, kz, launches};
                            b.build(c);
                            (void)run_timed(b, 1);  // warm-up
                            const double ms = run_timed(b, reps);
                            if (ms < best_ms) {
                                best_ms = ms;
                                best = b.cfg;
                            }
                        }
                }
        b.build(best);
        if (chosen) *chosen = b.cfg;
    });
}

bpx_status bpx_batch_execute(bpx_batch batch, void* hip_stream) {
    return batch_guard([&] {
        if (!batch) fail(BPX_ERR_INVALID_ARGUMENT, "NULL batch");
        BatchImpl& b = *batch->impl;
        const hipStream_t s = hip_stream ? (hipStream_t)hip_stream : b.stream;
        // one set of buffers: after the work on the batch's stream and the previous execution
        if (s != b.stream) {
            hip_check(hipEventRecord(b.ready, b.stream), "hipEventRecord");
            hip_check(hipStreamWaitEvent(s, b.ready, 0), "hipStreamWaitEvent");
        }
        hip_check(hipStreamWaitEvent(s, b.done, 0), "hipStreamWaitEvent");
        hip_check(hipGraphLaunch(b.exec, s), "hipGraphLaunch");
        hip_check(hipEventRecord(b.done, s), "hipEventRecord");
    });
}

bpx_status bpx_batch_synchronize(bpx_batch batch) {
    return batch_guard([&] {
        if (!batch) fail(BPX_ERR_INVALID_ARGUMENT, "NULL batch");
        batch->impl->wait();
    });
}

bpx_status bpx_batch_read_volume(bpx_batch batch, uint32_t index, float* voxel_volume) {
    return batch_guard([&] {
        if (!batch || !voxel_volume) fail(BPX_ERR_INVALID_ARGUMENT, "NULL argument");
        BatchImpl& b = *batch->impl;
        if (index >= b.plans.size()) fail(BPX_ERR_INVALID_ARGUMENT, "capture index out of range");
        b.wait();
        const PlanImpl& pl = *b.plans[index];
        const size_t nvox = (size_t)pl.g.n[0] * pl.g.n[1] * pl.g.n[2];
        hip_check(hipMemcpy(voxel_volume, pl.vol.ptr, (b.cplx ? 8u : 4u) * nvox, hipMemcpyDeviceToHost), "hipMemcpy");
    });
}

bpx_status bpx_batch_latencies(bpx_batch batch, double* ms) {
    return batch_guard([&] {
        if (!batch || !ms) fail(BPX_ERR_INVALID_ARGUMENT, "NULL argument");
        BatchImpl& b = *batch->impl;
        b.wait();
        const size_t n = b.plans.size();
        std::vector<uint64_t> st(n + 1);
        hip_check(hipMemcpy(st.data(), b.stamps.ptr, 8u * (n + 1), hipMemcpyDeviceToHost), "hipMemcpy");
        for (size_t i = 0; i < n; ++i)
            ms[i] = b.plans[i]->P ? (double)(int64_t)(st[1 + i] - st[0]) / b.wall_rate_khz : 0.0;
    });
}

bpx_status bpx_batch_benchmark(bpx_batch batch, int32_t warmup, int32_t iterations, bpx_batch_timing* out) {
    return batch_guard([&] {
        if (!batch || !out) fail(BPX_ERR_INVALID_ARGUMENT, "NULL argument");
        BatchImpl& b = *batch->impl;
        b.wait();
        if (warmup > 0) (void)run_timed(b, warmup);
        const int32_t it = std::max(iterations, 1);
        out->per_exec_ms = run_timed(b, it);
        out->iterations = it;
        uint64_t total = 0;
        for (uint64_t u : b.updates) total += u;
        out->updates = total;
        out->updates_per_s = out->per_exec_ms > 0.0 ? (double)total / (out->per_exec_ms * 1e-3) : 0.0;
    });
}

uint32_t bpx_batch_size(bpx_batch batch) { return batch ? (uint32_t)batch->impl->plans.size() : 0u; }

}  // extern "C"
