// Proposal net: parameters, forward / backward over dense level grids, Adam, octree feature
// and label assembly, training on (octree bounds, dense reconstruction) pairs, and the
// proposal hook used by the octree's child selection.  No reference counterpart.
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "bpx_host.hpp"

namespace bpx {

NetImpl::~NetImpl() {
    if (stream) (void)hipStreamDestroy(stream);
}

void NetImpl::ensure(uint64_t voxels) {
    if (voxels <= capacity) return;
    for (auto& a : act) a.alloc(voxels * 16u * 2u);
    logits.alloc(voxels * 16u * 4u);
    dlog.alloc(voxels * 16u * 2u);
    dx.alloc(voxels * 16u * 4u);
    dz.alloc(voxels * 16u * 2u);
    labels.alloc(voxels * 8u);
    mask.alloc(voxels);
    capacity = voxels;
}

void NetImpl::refresh_bf16(hipStream_t s) {
    hip_check(launch_to_bf16(static_cast<const float*>(params.ptr), static_cast<uint16_t*>(wb.ptr),
                             (uint64_t)kLayers * kW, s),
              "bf16 weights");
    for (int k = 0; k < kLayers; ++k)
        hip_check(launch_flip_transpose(static_cast<const uint16_t*>(wb.ptr) + (size_t)k * kW,
                                        static_cast<uint16_t*>(wt.ptr) + (size_t)k * kW, s),
                  "flipped weights");
}

static const float* bias_ptr(const NetImpl& n, int k) {
    return static_cast<const float*>(n.params.ptr) + (size_t)NetImpl::kLayers * NetImpl::kW + 16u * k;
}

void NetImpl::forward(uint32_t n0, uint32_t n1, uint32_t n2, hipStream_t s) {
    for (int k = 0; k < kLayers; ++k) {
        const uint16_t* W = static_cast<const uint16_t*>(wb.ptr) + (size_t)k * kW;
        const bool last = k == kLayers - 1;
        hip_check(launch_conv3_fwd(static_cast<const uint16_t*>(act[k].ptr), W, bias_ptr(*this, k),
                                   last ? nullptr : static_cast<uint16_t*>(act[k + 1].ptr),
                                   last ? static_cast<float*>(logits.ptr) : nullptr, n0, n1, n2, last ? 0 : 1,
                                   last ? kOut : 16u, s),
                  "conv forward");
    }
}

void NetImpl::backward(uint32_t n0, uint32_t n1, uint32_t n2, hipStream_t s) {
    const uint64_t V = (uint64_t)n0 * n1 * n2;
    uint16_t* cur = static_cast<uint16_t*>(dlog.ptr);  // dZ of the last layer
    uint16_t* other = static_cast<uint16_t*>(dz.ptr);
    float* g = static_cast<float*>(grads.ptr);
    for (int k = kLayers - 1; k >= 0; --k) {
        hip_check(launch_conv3_dw(cur, static_cast<const uint16_t*>(act[k].ptr), g + (size_t)k * kW,
                                  g + (size_t)kLayers * kW + 16u * k, n0, n1, n2, s),
                  "conv weight gradient");
        if (k == 0) break;
        hip_check(launch_conv3_fwd(cur, static_cast<const uint16_t*>(wt.ptr) + (size_t)k * kW, nullptr, nullptr,
                                   static_cast<float*>(dx.ptr), n0, n1, n2, 0, 16u, s),
                  "conv data gradient");
        hip_check(launch_relu_back(static_cast<const float*>(dx.ptr), static_cast<const uint16_t*>(act[k].ptr),
                                   other, V * 16u, s),
                  "relu backward");
        std::swap(cur, other);
    }
}

void NetImpl::adam(float lr, hipStream_t s) {
    ++step;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)step), c2 = 1.0f - std::pow(b2, (float)step);
    hip_check(launch_adam(static_cast<float*>(params.ptr), static_cast<const float*>(grads.ptr),
                          static_cast<float*>(adam_m.ptr), static_cast<float*>(adam_v.ptr), nullptr, kParams, 0,
                          lr, b1, b2, 1e-8f, c1, c2, s),
              "adam");
    refresh_bf16(s);
}

static NetImpl* make_net(uint64_t seed) {
    if (!device_ready()) fail(BPX_ERR_NO_DEVICE, "no gfx950 device available");
    auto n = std::make_unique<NetImpl>();
    hip_check(hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking), "hipStreamCreate");
    n->params.alloc(4ull * NetImpl::kParams);
    n->grads.alloc(4ull * NetImpl::kParams);
    n->adam_m.alloc(4ull * NetImpl::kParams);
    n->adam_v.alloc(4ull * NetImpl::kParams);
    n->wb.alloc(2ull * NetImpl::kLayers * NetImpl::kW);
    n->wt.alloc(2ull * NetImpl::kLayers * NetImpl::kW);
    n->loss.alloc(sizeof(double));
    n->vmax.alloc(sizeof(double));
    n->vmax_part.alloc(sizeof(double) * NetImpl::kMaxParts);
    // He initialisation over the real fan-in; padded input (layer 0) and output (layer 3)
    // channels start at zero; biases zero.
    std::vector<float> p(NetImpl::kParams, 0.0f);
    std::mt19937_64 rng(seed);
    for (int k = 0; k < NetImpl::kLayers; ++k) {
        const uint32_t cin = k == 0 ? NetImpl::kIn : 16u, cout = k == NetImpl::kLayers - 1 ? NetImpl::kOut : 16u;
        std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (27.0f * cin)));
        for (uint32_t t = 0; t < 27; ++t)
            for (uint32_t o = 0; o < cout; ++o)
                for (uint32_t i = 0; i < cin; ++i) p[(size_t)k * NetImpl::kW + (t * 16u + o) * 16u + i] = nd(rng);
    }
    hip_check(hipMemcpy(n->params.ptr, p.data(), 4ull * p.size(), hipMemcpyHostToDevice), "hipMemcpy");
    hip_check(hipMemset(n->adam_m.ptr, 0, 4ull * NetImpl::kParams), "hipMemset");
    hip_check(hipMemset(n->adam_v.ptr, 0, 4ull * NetImpl::kParams), "hipMemset");
    n->refresh_bf16(n->stream);
    hip_check(hipStreamSynchronize(n->stream), "hipStreamSynchronize");
    return n.release();
}

// Features of level l (dense grid) into the net's input buffer; with_mask also marks the
// evaluated cells in net.mask.
static void level_features(NetImpl& net, const OctreeState::Level& lv, bool with_mask, hipStream_t s) {
    const uint64_t V = (uint64_t)lv.n[0] * lv.n[1] * lv.n[2];
    net.ensure(V);
    hip_check(hipMemsetAsync(net.act[0].ptr, 0, V * 32u, s), "hipMemsetAsync");
    if (with_mask) hip_check(hipMemsetAsync(net.mask.ptr, 0, V, s), "hipMemsetAsync");
    const float level_code = std::log2((float)lv.B) / 8.0f;
    hip_check(launch_oct_features(static_cast<const uint32_t*>(lv.list.ptr), static_cast<const float*>(lv.U.ptr),
                                  static_cast<const float*>(lv.V.ptr), lv.count,
                                  static_cast<const double*>(lv.stats.ptr), lv.n[1], lv.n[2], level_code,
                                  static_cast<uint16_t*>(net.act[0].ptr),
                                  with_mask ? static_cast<uint8_t*>(net.mask.ptr) : nullptr, s),
              "octree features");
}

bool net_child_flags(PlanImpl& pl, uint32_t l, hipStream_t s) {
    OctreeState& st = *pl.oct;
    if (!st.net) return false;
    NetImpl& net = *st.net;
    const OctreeState::Level& lv = st.levels[l];
    level_features(net, lv, false, s);
    net.forward(lv.n[0], lv.n[1], lv.n[2], s);
    const float p = st.prob[l];
    const float logit_thr = std::log(p / (1.0f - p));
    // low-confidence band (fallback probability, p]: full-resolution fallback; none when unset
    const float fb = st.fallback[l];
    const float fallback_thr = fb > 0.0f && fb < p ? std::log(fb / (1.0f - fb)) : logit_thr;
    hip_check(launch_oct_child_flags(static_cast<const uint32_t*>(lv.list.ptr), static_cast<const float*>(lv.U.ptr),
                                     lv.count, static_cast<const double*>(st.stats.ptr),
                                     static_cast<const float*>(net.logits.ptr), lv.n[1], lv.n[2], logit_thr,
                                     fallback_thr, st.keep, static_cast<uint32_t*>(st.flags.ptr), s),
              "learned child flags");
    return true;
}

}  // namespace bpx

using namespace bpx;

struct bpx_net_s {
    std::unique_ptr<NetImpl> impl;
};

namespace {
template <class F>
bpx_status net_guard(F&& f) {
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

void check_dims(const uint32_t dims[3]) {
    if (!dims || !dims[0] || !dims[1] || !dims[2]) fail(BPX_ERR_INVALID_ARGUMENT, "grid dims must be > 0");
    if ((uint64_t)dims[0] * dims[1] * dims[2] > (1ull << 26)) fail(BPX_ERR_UNSUPPORTED, "grid too large");
}

void upload_input(NetImpl& n, const float* x, uint64_t V, hipStream_t s) {
    DeviceBuffer tmp;
    tmp.alloc(V * 16u * 4u);
    hip_check(hipMemcpyAsync(tmp.ptr, x, V * 64u, hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    hip_check(launch_to_bf16(static_cast<const float*>(tmp.ptr), static_cast<uint16_t*>(n.act[0].ptr), V * 16u, s),
              "bf16 input");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}
}  // namespace

extern "C" {

uint32_t bpx_net_param_count(void) { return NetImpl::kParams; }

bpx_status bpx_net_create(uint64_t seed, bpx_net* out) {
    return net_guard([&] {
        if (!out) fail(BPX_ERR_INVALID_ARGUMENT, "NULL output");
        *out = nullptr;
        auto h = std::make_unique<bpx_net_s>();
        h->impl.reset(make_net(seed));
        *out = h.release();
    });
}

void bpx_net_destroy(bpx_net net) {
    if (!net) return;
    (void)hipStreamSynchronize(net->impl->stream);
    delete net;
}

bpx_status bpx_net_get_params(bpx_net net, float* params) {
    return net_guard([&] {
        if (!net || !params) fail(BPX_ERR_INVALID_ARGUMENT, "NULL argument");
        hip_check(hipStreamSynchronize(net->impl->stream), "hipStreamSynchronize");
        hip_check(hipMemcpy(params, net->impl->params.ptr, 4ull * NetImpl::kParams, hipMemcpyDeviceToHost), "hipMemcpy");
    });
}

bpx_status bpx_net_set_params(bpx_net net, const float* params) {
    return net_guard([&] {
        if (!net || !params) fail(BPX_ERR_INVALID_ARGUMENT, "NULL argument");
        for (uint32_t i = 0; i < NetImpl::kParams; ++i)
            if (!std::isfinite(params[i])) fail(BPX_ERR_INVALID_ARGUMENT, "non-finite parameter");
        NetImpl& n = *net->impl;
        hip_check(hipMemcpy(n.params.ptr, params, 4ull * NetImpl::kParams, hipMemcpyHostToDevice), "hipMemcpy");
        hip_check(hipMemset(n.adam_m.ptr, 0, 4ull * NetImpl::kParams), "hipMemset");
        hip_check(hipMemset(n.adam_v.ptr, 0, 4ull * NetImpl::kParams), "hipMemset");
        n.step = 0;
        n.refresh_bf16(n.stream);
        hip_check(hipStreamSynchronize(n.stream), "hipStreamSynchronize");
    });
}

bpx_status bpx_net_forward(bpx_net net, const float* input, const uint32_t dims[3], float* logits) {
    return net_guard([&] {
        if (!net || !input || !logits) fail(BPX_ERR_INVALID_ARGUMENT, "NULL argument");
        check_dims(dims);
        NetImpl& n = *net->impl;
        const uint64_t V = (uint64_t)dims[0] * dims[1] * dims[2];
        n.ensure(V);
        upload_input(n, input, V, n.stream);
        n.forward(dims[0], dims[1], dims[2], n.stream);
        hip_check(hipMemcpyAsync(logits, n.logits.ptr, V * 64u, hipMemcpyDeviceToHost, n.stream), "hipMemcpyAsync");
        hip_check(hipStreamSynchronize(n.stream), "hipStreamSynchronize");
    });
}

bpx_status bpx_net_gradients(bpx_net net, const float* input, const uint32_t dims[3], const uint8_t* labels,
                             const uint8_t* mask, float pos_weight, double* loss, float* grads) {
    return net_guard([&] {
        if (!net || !input || !labels || !loss || !grads) fail(BPX_ERR_INVALID_ARGUMENT, "NULL argument");
        check_dims(dims);
        NetImpl& n = *net->impl;
        const hipStream_t s = n.stream;
        const uint64_t V = (uint64_t)dims[0] * dims[1] * dims[2];
        n.ensure(V);
        upload_input(n, input, V, s);
        DeviceBuffer dmask;
        if (mask) {
            dmask.alloc(V);
            hip_check(hipMemcpy(dmask.ptr, mask, V, hipMemcpyHostToDevice), "hipMemcpy");
        }
        hip_check(hipMemcpy(n.labels.ptr, labels, V * 8u, hipMemcpyHostToDevice), "hipMemcpy");
        hip_check(hipMemsetAsync(n.grads.ptr, 0, 4ull * NetImpl::kParams, s), "hipMemsetAsync");
        hip_check(hipMemsetAsync(n.loss.ptr, 0, sizeof(double), s), "hipMemsetAsync");
        n.forward(dims[0], dims[1], dims[2], s);
        hip_check(launch_bce(static_cast<const float*>(n.logits.ptr), static_cast<const uint8_t*>(n.labels.ptr),
                             mask ? static_cast<const uint8_t*>(dmask.ptr) : nullptr, (uint32_t)V, NetImpl::kOut,
                             pos_weight, 1.0f, static_cast<uint16_t*>(n.dlog.ptr), static_cast<double*>(n.loss.ptr), s),
                  "bce");
        n.backward(dims[0], dims[1], dims[2], s);
        hip_check(hipMemcpyAsync(grads, n.grads.ptr, 4ull * NetImpl::kParams, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
        hip_check(hipMemcpyAsync(loss, n.loss.ptr, sizeof(double), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    });
}

bpx_status bpx_net_train_octree(bpx_net net, bpx_plan octree_plan, bpx_plan dense_plan,
                                float occupied_fraction, float pos_weight, float lr, double* loss) {
    return net_guard([&] {
        if (!net || !octree_plan || !dense_plan || !loss) fail(BPX_ERR_INVALID_ARGUMENT, "NULL argument");
        PlanImpl& op = *octree_plan->impl;
        PlanImpl& dp = *dense_plan->impl;
        if (!op.oct) fail(BPX_ERR_INVALID_ARGUMENT, "octree_plan is not an octree plan");
        if (std::memcmp(op.g.n, dp.g.n, sizeof(op.g.n)) != 0)
            fail(BPX_ERR_INVALID_ARGUMENT, "octree and dense plans have different grids");
        OctreeState& st = *op.oct;
        if (st.levels[0].count == 0) fail(BPX_ERR_INVALID_ARGUMENT, "the octree plan has not been executed");
        NetImpl& n = *net->impl;
        const hipStream_t s = n.stream;
        hip_check(hipStreamSynchronize(op.stream), "hipStreamSynchronize");
        hip_check(hipStreamSynchronize(dp.stream), "hipStreamSynchronize");
        const uint64_t nvox = (uint64_t)dp.g.n[0] * dp.g.n[1] * dp.g.n[2];
        // max-pooled pyramid of the dense volume: pooled[l] at level-l resolution, l = 1..L
        std::vector<uint64_t> off(st.L + 1, 0);
        uint64_t total = 0;
        for (uint32_t l = 1; l < st.L; ++l) {
            off[l] = total;
            total += (uint64_t)st.levels[l].n[0] * st.levels[l].n[1] * st.levels[l].n[2];
        }
        if (n.pooled_a.bytes < 4 * std::max<uint64_t>(total, 1)) n.pooled_a.alloc(4 * std::max<uint64_t>(total, 1));
        hip_check(launch_max_f64(static_cast<const float*>(dp.vol.ptr), nvox, static_cast<double*>(n.vmax_part.ptr),
                                 NetImpl::kMaxParts, static_cast<double*>(n.vmax.ptr), s),
                  "FP64 max");
        const float* finer = static_cast<const float*>(dp.vol.ptr);
        for (int l = (int)st.L - 1; l >= 1; --l) {
            const auto& lv = st.levels[l];
            float* dst = static_cast<float*>(n.pooled_a.ptr) + off[l];
            hip_check(launch_maxpool2(finer, dst, lv.n[0], lv.n[1], lv.n[2], s), "max pool");
            finer = dst;
        }
        hip_check(hipMemsetAsync(n.grads.ptr, 0, 4ull * NetImpl::kParams, s), "hipMemsetAsync");
        hip_check(hipMemsetAsync(n.loss.ptr, 0, sizeof(double), s), "hipMemsetAsync");
        for (uint32_t l = 0; l < st.L; ++l) {
            const auto& lv = st.levels[l];
            if (lv.count == 0) break;
            const uint64_t V = (uint64_t)lv.n[0] * lv.n[1] * lv.n[2];
            level_features(n, lv, true, s);
            const float* children = l + 1 < st.L ? static_cast<const float*>(n.pooled_a.ptr) + off[l + 1]
                                                 : static_cast<const float*>(dp.vol.ptr);
            hip_check(launch_child_labels(children, lv.n[0], lv.n[1], lv.n[2], static_cast<const double*>(n.vmax.ptr),
                                          occupied_fraction, static_cast<uint8_t*>(n.labels.ptr), s),
                      "child labels");
            n.forward(lv.n[0], lv.n[1], lv.n[2], s);
            // every level weighs equally: mean over its evaluated blocks x 8 children
            const float scale = 1.0f / (float)((uint64_t)lv.count * NetImpl::kOut * st.L);
            hip_check(launch_bce(static_cast<const float*>(n.logits.ptr), static_cast<const uint8_t*>(n.labels.ptr),
                                 static_cast<const uint8_t*>(n.mask.ptr), (uint32_t)V, NetImpl::kOut, pos_weight, scale,
                                 static_cast<uint16_t*>(n.dlog.ptr), static_cast<double*>(n.loss.ptr), s),
                      "bce");
            n.backward(lv.n[0], lv.n[1], lv.n[2], s);
        }
        n.adam(lr, s);
        hip_check(hipMemcpyAsync(loss, n.loss.ptr, sizeof(double), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    });
}

bpx_status bpx_octree_plan_attach_net(bpx_plan plan, bpx_net net) {
    return net_guard([&] {
        if (!plan) fail(BPX_ERR_INVALID_ARGUMENT, "NULL plan");
        if (!plan->impl->oct) fail(BPX_ERR_INVALID_ARGUMENT, "not an octree plan");
        plan->impl->oct->net = net ? net->impl.get() : nullptr;
    });
}

}  // extern "C"
