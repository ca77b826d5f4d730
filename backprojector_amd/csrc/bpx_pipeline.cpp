// Host pipeline of the reference (bp::backproject, bp::phasor_reconstruction), restated in
// C++ over plain arrays, driving the device plans.  Float arithmetic follows the reference's
// expression order; this file is compiled with -ffp-contract=off.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>

#include "bpx_host.hpp"

namespace bpx {
namespace pipeline {

// backproject.hpp:42-53: float differences, double squares/sum/sqrt.
static double ref_distance(const float* p1, const float* p2) {
    double t[3];
    for (int i = 0; i < 3; ++i) {
        t[i] = (double)(p1[i] - p2[i]);
        t[i] = t[i] * t[i];
    }
    return std::sqrt(t[0] + t[1] + t[2]);
}

// Reverses the axes of a C-order array (a column-major array read as C order is the
// transpose of the row-major one).
static std::vector<float> reverse_axes(const float* a, const std::vector<int64_t>& shape_in,
                                       std::vector<int64_t>& shape_out) {
    const int nd = (int)shape_in.size();
    shape_out.assign(shape_in.rbegin(), shape_in.rend());
    size_t n = 1;
    for (auto s : shape_in) n *= (size_t)s;
    std::vector<float> out(n);
    std::vector<int64_t> stride_in(nd, 1);
    for (int d = nd - 2; d >= 0; --d) stride_in[d] = stride_in[d + 1] * shape_in[d + 1];
    std::vector<int64_t> idx(nd, 0);  // index into shape_out
    for (size_t o = 0; o < n; ++o) {
        int64_t off = 0;
        for (int d = 0; d < nd; ++d) off += idx[d] * stride_in[nd - 1 - d];
        out[o] = a[off];
        for (int d = nd - 1; d >= 0; --d) {
            if (++idx[d] < shape_out[d]) break;
            idx[d] = 0;
        }
    }
    return out;
}

Arrays normalise(const float* data, const int64_t* data_shape, int data_ndim, const float* cgrid,
                 const int64_t* cgrid_shape, const float* lgrid, const int64_t* lgrid_shape,
                 bool row_major, bool confocal) {
    if (!data || !data_shape || !cgrid || !cgrid_shape || !lgrid || !lgrid_shape)
        fail(BPX_ERR_INVALID_ARGUMENT, "NULL array argument");
    const int want = confocal ? 3 : 5;
    if (data_ndim != want)
        fail(BPX_ERR_INVALID_ARGUMENT, std::string("transient_data must have ") +
                                           (confocal ? "3 dims [X][Y][T]" : "5 dims [Lx][Ly][Cx][Cy][T]"));
    Arrays a;
    std::vector<int64_t> ds(data_shape, data_shape + data_ndim), cs(cgrid_shape, cgrid_shape + 3),
        ls(lgrid_shape, lgrid_shape + 3);
    for (auto v : ds)
        if (v <= 0) fail(BPX_ERR_INVALID_ARGUMENT, "empty transient_data dimension");
    if (row_major) {
        a.data_p = data;
        a.cgrid_p = cgrid;
        a.lgrid_p = lgrid;
        a.data_shape = ds;
        a.cgrid_shape = cs;
        a.lgrid_shape = ls;
    } else {
        a.data = reverse_axes(data, ds, a.data_shape);
        a.cgrid = reverse_axes(cgrid, cs, a.cgrid_shape);
        a.lgrid = reverse_axes(lgrid, ls, a.lgrid_shape);
        a.data_p = a.data.data();
        a.cgrid_p = a.cgrid.data();
        a.lgrid_p = a.lgrid.data();
    }
    if (a.cgrid_shape[2] != 3 || a.lgrid_shape[2] != 3)
        fail(BPX_ERR_INVALID_ARGUMENT, "grid positions must be [X][Y][3] (row-major)");
    if (confocal) {
        if (a.data_shape[0] != a.cgrid_shape[0] || a.data_shape[1] != a.cgrid_shape[1])
            fail(BPX_ERR_INVALID_ARGUMENT, "confocal transient_data [X][Y][T] must match the camera grid");
    } else {
        if (a.data_shape[0] != a.lgrid_shape[0] || a.data_shape[1] != a.lgrid_shape[1] ||
            a.data_shape[2] != a.cgrid_shape[0] || a.data_shape[3] != a.cgrid_shape[1])
            fail(BPX_ERR_INVALID_ARGUMENT,
                 "exhaustive transient_data [Lx][Ly][Cx][Cy][T] must match the laser and camera grids");
    }
    return a;
}

std::vector<bpx_pair> precompute_distances(const float laser_position[3], const float* lgrid,
                                           uint32_t Lx, uint32_t Ly,
                                           const float camera_position[3], const float* cgrid,
                                           uint32_t Cx, uint32_t Cy, bool confocal) {
    std::vector<bpx_pair> out;
    if (confocal) {
        // :399-413 -- both legs from the camera grid and camera_position
        out.resize((size_t)Cx * Cy);
        for (uint32_t x = 0; x < Cx; ++x)
            for (uint32_t y = 0; y < Cy; ++y) {
                const float* cp = cgrid + 3 * ((size_t)x * Cy + y);
                bpx_pair& pr = out[(size_t)x * Cy + y];
                std::memcpy(pr.laser_point, cp, 12);
                std::memcpy(pr.camera_point, cp, 12);
                const float w = (float)ref_distance(camera_position, cp);
                pr.camera_wall = w;
                pr.laser_wall = w;
            }
        return out;
    }
    // :374-397
    out.resize((size_t)Lx * Ly * Cx * Cy);
    for (uint32_t lx = 0; lx < Lx; ++lx)
        for (uint32_t ly = 0; ly < Ly; ++ly) {
            const float* lp = lgrid + 3 * ((size_t)lx * Ly + ly);
            const float lw = (float)ref_distance(laser_position, lp);
            for (uint32_t cx = 0; cx < Cx; ++cx)
                for (uint32_t cy = 0; cy < Cy; ++cy) {
                    const size_t p = (((size_t)lx * Ly + ly) * Cx + cx) * Cy + cy;
                    const float* cp = cgrid + 3 * ((size_t)cx * Cy + cy);
                    std::memcpy(out[p].laser_point, lp, 12);
                    std::memcpy(out[p].camera_point, cp, 12);
                    out[p].camera_wall = (float)ref_distance(camera_position, cp);
                    out[p].laser_wall = lw;
                }
        }
    return out;
}

static void min_max(const float* a, int na, const float* b, int nb, float* mn, float* mx) {
    float lo = FLT_MAX, hi = 0.0f;
    for (int i = 0; i < na; ++i)
        for (int j = 0; j < nb; ++j) {
            const float d = (float)ref_distance(a + 3 * i, b + 3 * j);
            lo = std::min(d, lo);
            hi = std::max(d, hi);
        }
    *mn = lo;
    *mx = hi;
}

void time_limits(const float laser_position[3], const float* lgrid, uint32_t Lx, uint32_t Ly,
                 const float camera_position[3], const float* cgrid, uint32_t Cx, uint32_t Cy,
                 const float volume_size[3], const float volume_position[3], float t0, float dT,
                 float* min_idx, float* max_idx) {
    float las[4][3], cam[4][3], vol[8][3];
    const uint32_t lxs[2] = {0, Lx - 1}, lys[2] = {0, Ly - 1};
    for (int m = 0; m < 4; ++m)
        std::memcpy(las[m], lgrid + 3 * ((size_t)lxs[(m >> 1) & 1] * Ly + lys[m & 1]), 12);
    if (Cx >= 3) {
        // :651-652 -- the row-major camera grid is transposed to [3][Y][X] before its corners
        // are taken; distance() then reads the first three entries of each length-X corner row.
        const uint32_t cs[2] = {0, 2}, cys[2] = {0, Cy - 1};
        for (int m = 0; m < 4; ++m)
            for (int k = 0; k < 3; ++k)
                cam[m][k] = cgrid[3 * ((size_t)k * Cy + cys[m & 1]) + cs[(m >> 1) & 1]];
    } else {
        // fewer than 3 rows: the reference reads past the array; use the grid's own corners
        const uint32_t cxs[2] = {0, Cx - 1}, cys[2] = {0, Cy - 1};
        for (int m = 0; m < 4; ++m)
            std::memcpy(cam[m], cgrid + 3 * ((size_t)cxs[(m >> 1) & 1] * Cy + cys[m & 1]), 12);
    }
    for (int m = 0; m < 8; ++m)
        for (int k = 0; k < 3; ++k) {
            const float lower = volume_position[k] - volume_size[k] / 2.0f;
            const float mult = (m & (1 << (2 - k))) ? 1.0f : 0.0f;
            vol[m][k] = lower + mult * volume_size[k];
        }
    float a, b, c, d, e, f, g, h;
    // origins are single points (the reference's 1-D origin indexing is undefined)
    min_max(laser_position, 1, &las[0][0], 4, &a, &b);
    min_max(camera_position, 1, &cam[0][0], 4, &c, &d);
    min_max(&vol[0][0], 8, &las[0][0], 4, &e, &f);
    min_max(&vol[0][0], 8, &cam[0][0], 4, &g, &h);
    float mn = a + c;
    mn = mn + e;
    mn = mn + g;
    float mx = b + d;
    mx = mx + f;
    mx = mx + h;
    mn = mn - t0;
    mx = mx - t0;
    *min_idx = mn / dT - 100.0f;
    *max_idx = mx / dT + 100.0f;
}

// Right-handed wall-aligned axes centred on the volume position (BPX_FLAG_CENTRED_WALL_BASIS):
// y = -normal, x = unit(y x helper) with helper world z (world x when |y . z| > 0.9), z = x x y;
// axis r advances volume_size[r] / (N[r] - 1) per voxel, and the volume's middle is at
// volume_position.
static void centred_basis(const float volume_position[3], const float volume_size[3], const uint32_t n[3],
                          const float wall_normal[3], float zero[3], float inc[9]) {
    const double nn = std::sqrt((double)wall_normal[0] * wall_normal[0] + (double)wall_normal[1] * wall_normal[1] +
                                (double)wall_normal[2] * wall_normal[2]);
    if (!(nn > 0.0) || !std::isfinite(nn)) fail(BPX_ERR_INVALID_ARGUMENT, "wall_normal must be finite");
    const double y[3] = {-wall_normal[0] / nn, -wall_normal[1] / nn, -wall_normal[2] / nn};
    const double h[3] = {std::fabs(y[2]) > 0.9 ? 1.0 : 0.0, 0.0, std::fabs(y[2]) > 0.9 ? 0.0 : 1.0};
    double x[3] = {y[1] * h[2] - y[2] * h[1], y[2] * h[0] - y[0] * h[2], y[0] * h[1] - y[1] * h[0]};
    const double xn = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    for (double& c : x) c /= xn;
    const double z[3] = {x[1] * y[2] - x[2] * y[1], x[2] * y[0] - x[0] * y[2], x[0] * y[1] - x[1] * y[0]};
    const double* ax[3] = {x, y, z};
    for (int k = 0; k < 3; ++k) {
        double half = 0.0;
        for (int r = 0; r < 3; ++r) half += ax[r][k] * (double)volume_size[r];
        zero[k] = (float)((double)volume_position[k] - 0.5 * half);
    }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) inc[3 * r + c] = (float)(ax[r][c] * (double)volume_size[r]) / (float)(n[r] - 1u);
}

void basis(const float volume_position[3], const float volume_size[3], const uint32_t n[3],
           const float wall_normal[3], float zero[3], float inc[9], bool centred) {
    if (centred && (wall_normal[0] != 0.0f || wall_normal[1] != 0.0f || wall_normal[2] != 0.0f)) {
        centred_basis(volume_position, volume_size, n, wall_normal, zero, inc);
        return;
    }
    float ax[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    if (wall_normal[0] != 0.0f || wall_normal[1] != 0.0f || wall_normal[2] != 0.0f) {
        const float y[3] = {-wall_normal[0], -wall_normal[1], -wall_normal[2]};
        const float z[3] = {0, 0, 1};
        const float x[3] = {y[1] * z[2] - y[2] * z[1], y[2] * z[0] - y[0] * z[2], y[0] * z[1] - y[1] * z[0]};
        const float z2[3] = {y[1] * x[2] - y[2] * x[1], y[2] * x[0] - y[0] * x[2], y[0] * x[1] - y[1] * x[0]};
        const float* v3[3] = {x, y, z2};
        for (int r = 0; r < 3; ++r) {
            const float* v = v3[r];
            const float aa = v[0] * v[0], bb = v[1] * v[1], cc = v[2] * v[2];
            const double nrm = std::sqrt(((double)aa + (double)bb) + (double)cc);
            for (int c = 0; c < 3; ++c) ax[r][c] = (float)((double)v[c] / nrm);
        }
    }
    for (int k = 0; k < 3; ++k) zero[k] = volume_position[k] - volume_size[k] / 2.0f;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) inc[3 * r + c] = (ax[r][c] * volume_size[c]) / (float)(n[c] - 1u);
}

ChunkWindow chunk_window(uint32_t T, float min_idx, float max_idx) {
    const float fmin = std::floor(min_idx), fmax = std::ceil(max_idx);
    if (!(fmax - fmin >= 1.0f) || fmax - fmin > 4194304.0f)
        fail(BPX_ERR_INVALID_ARGUMENT, "time limits give an empty or oversized chunk");
    ChunkWindow w{};
    w.Tc = (uint32_t)(fmax - fmin);
    const uint64_t from = fmin > 0.0f ? (uint64_t)fmin : 0;
    uint64_t to = (uint64_t)(fmax > 0.0f ? fmax : 0.0f);
    if (to > T) to = T;
    const uint64_t cfrom = fmin > 0.0f ? 0 : (uint64_t)std::fabs(fmin);
    // backproject.hpp:701-706; a row window that would overrun the chunk is not copied
    if (to > from && cfrom + (to - from) <= w.Tc) {
        w.from = (uint32_t)from;
        w.cfrom = (uint32_t)cfrom;
        w.ncopy = (uint32_t)(to - from);
    }
    return w;
}

std::vector<float> transient_chunk(const float* data, uint64_t rows, uint32_t T, float min_idx,
                                   float max_idx, uint32_t* Tc_out) {
    const ChunkWindow w = chunk_window(T, min_idx, max_idx);
    std::vector<float> out(rows * w.Tc, 0.0f);
    if (w.ncopy)
        for (uint64_t r = 0; r < rows; ++r)
            std::memcpy(&out[r * w.Tc + w.cfrom], data + r * T + w.from, 4 * (size_t)w.ncopy);
    *Tc_out = w.Tc;
    return out;
}

int phasor_taps(float wavelength, float dT, uint32_t times, float sigma, std::vector<float>& c,
                std::vector<float>& s) {
    c.clear();
    s.clear();
    if (!(sigma > 0.0f)) {
        // backproject.hpp:434-451
        const int pulse = (int)(((float)times * wavelength) / dT);
        const int cycle = (int)(wavelength / dT);
        if (pulse <= 0 || cycle <= 0 || pulse > 65536) return 0;
        const float vsigma = ((float)times * wavelength) / 6.0f;
        const float denom = (2.0f * vsigma) * vsigma;
        for (int k = 1; k <= pulse; ++k) {
            const double t = (double)dT * ((double)k - pulse / 2.0);
            const double gk = std::exp(-t * t / (double)denom);
            const double prog = 2.0 * M_PI * ((1.0 / cycle) * (double)k);
            const float sn = (float)std::sin(prog), cs = (float)std::cos(prog);
            c.push_back((float)((double)cs * gk));
            s.push_back((float)((double)sn * gk));
        }
        return pulse;
    }
    // Extension: explicit sigma, window +-3 sigma, carrier of exactly `wavelength`.
    const int half = (int)std::ceil(3.0 * (double)sigma / (double)dT);
    const int pulse = 2 * half + 1;
    if (pulse > 65536) return 0;
    for (int k = 1; k <= pulse; ++k) {
        const double t = (double)dT * ((double)k - pulse / 2.0);
        const double gk = std::exp(-t * t / (2.0 * (double)sigma * (double)sigma));
        const double prog = 2.0 * M_PI * (double)k * (double)dT / (double)wavelength;
        c.push_back((float)(std::cos(prog) * gk));
        s.push_back((float)(std::sin(prog) * gk));
    }
    return pulse;
}

}  // namespace pipeline

static void upload_taps(const std::vector<float>& c, const std::vector<float>& s, DeviceBuffer& dc,
                        DeviceBuffer& ds) {
    if (c.empty() || c.size() > 4096 || c.size() != s.size())
        fail(BPX_ERR_UNSUPPORTED, "phasor pulses of 1..4096 taps are supported");
    dc.alloc(sizeof(float) * c.size());
    ds.alloc(sizeof(float) * s.size());
    hip_check(hipMemcpy(dc.ptr, c.data(), sizeof(float) * c.size(), hipMemcpyHostToDevice), "hipMemcpy");
    hip_check(hipMemcpy(ds.ptr, s.data(), sizeof(float) * s.size(), hipMemcpyHostToDevice), "hipMemcpy");
}

// Filters host rows on the device (phasor_fir_kernel) into host complex output [rows][T].
void phasor_filter_host(const float* in, uint64_t rows, uint32_t T, const std::vector<float>& c,
                        const std::vector<float>& s, float* out) {
    if (rows == 0 || T == 0) return;
    DeviceBuffer din, dout, dc, ds;
    upload_taps(c, s, dc, ds);
    if (rows * T >= (1ull << 38)) fail(BPX_ERR_UNSUPPORTED, "phasor input too large");
    din.alloc(sizeof(float) * rows * T);
    dout.alloc(sizeof(float) * 2 * rows * T);
    hip_check(hipMemcpy(din.ptr, in, sizeof(float) * rows * T, hipMemcpyHostToDevice), "hipMemcpy");
    hip_check(launch_phasor_fir(static_cast<const float*>(din.ptr), (uint32_t)rows, T,
                                static_cast<const float*>(dc.ptr), static_cast<const float*>(ds.ptr),
                                (int)c.size(), dout.ptr, T, 0u, kC64, 1.0f, nullptr),
              "phasor FIR kernel");
    hip_check(hipMemcpy(out, dout.ptr, sizeof(float) * 2 * rows * T, hipMemcpyDeviceToHost), "hipMemcpy(phasor)");
}

// The chunk on the device: the source rows in one copy, the chunk window cropped
// (get_transient_chunk) into [rows][Tc] real rows.
static DeviceBuffer chunk_rows_device(const float* src, uint64_t rows, uint32_t src_T, const pipeline::ChunkWindow& w) {
    if (rows >= (1ull << 32) || rows * w.Tc >= (1ull << 38)) fail(BPX_ERR_UNSUPPORTED, "phasor input too large");
    DeviceBuffer full, chunk;
    chunk.alloc(sizeof(float) * std::max<uint64_t>(rows * w.Tc, 1));
    if (rows == 0) return chunk;
    full.alloc(sizeof(float) * rows * src_T);
    upload_host(full.ptr, src, sizeof(float) * rows * src_T, nullptr);
    hip_check(launch_crop_rows(static_cast<const float*>(full.ptr), src_T, nullptr, (uint32_t)rows, w.Tc, w.from,
                               w.cfrom, w.ncopy, static_cast<float*>(chunk.ptr), nullptr),
              "crop kernel");
    hip_check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
    return chunk;
}

// The phasor FIR of device chunk rows into [rows][Tc] complex rows (no host round trip).
static DeviceBuffer phasor_rows_device(const DeviceBuffer& chunk, uint64_t rows, uint32_t Tc,
                                       const std::vector<float>& c, const std::vector<float>& s) {
    DeviceBuffer out, dc, ds;
    upload_taps(c, s, dc, ds);
    out.alloc(sizeof(float) * 2 * std::max<uint64_t>(rows * Tc, 1));
    if (rows == 0) return out;
    hip_check(launch_phasor_fir(static_cast<const float*>(chunk.ptr), (uint32_t)rows, Tc,
                                static_cast<const float*>(dc.ptr), static_cast<const float*>(ds.ptr),
                                (int)c.size(), out.ptr, Tc, 0u, kC64, 1.0f, nullptr),
              "phasor FIR kernel");
    hip_check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
    return out;
}

}  // namespace bpx

using namespace bpx;

namespace {

struct PipelineGeometry {
    pipeline::Arrays arr;
    uint32_t Lx, Ly, Cx, Cy, T;
    uint64_t rows;
    float min_idx, max_idx;
    std::vector<bpx_pair> pairs;
};

PipelineGeometry prepare(const float* transient_data, const int64_t* data_shape, int data_ndim,
                         const float* cgrid, const int64_t* cgrid_shape, const float* lgrid,
                         const int64_t* lgrid_shape, const float camera_position[3],
                         const float laser_position[3], float t0, float deltaT, int is_confocal,
                         const float volume_position[3], const float volume_size[3],
                         const uint32_t vps[3], int assume_row_major) {
    if (!camera_position || !laser_position || !volume_position || !volume_size || !vps)
        fail(BPX_ERR_INVALID_ARGUMENT, "NULL argument");
    PipelineGeometry g;
    g.arr = pipeline::normalise(transient_data, data_shape, data_ndim, cgrid, cgrid_shape, lgrid,
                                lgrid_shape, assume_row_major != 0, is_confocal != 0);
    g.Cx = (uint32_t)g.arr.cgrid_shape[0];
    g.Cy = (uint32_t)g.arr.cgrid_shape[1];
    g.Lx = (uint32_t)g.arr.lgrid_shape[0];
    g.Ly = (uint32_t)g.arr.lgrid_shape[1];
    g.T = (uint32_t)g.arr.data_shape.back();
    g.rows = 1;
    for (size_t d = 0; d + 1 < g.arr.data_shape.size(); ++d) g.rows *= (uint64_t)g.arr.data_shape[d];
    if (!(deltaT > 0.0f)) fail(BPX_ERR_INVALID_ARGUMENT, "deltaT must be > 0");
    pipeline::time_limits(laser_position, g.arr.lgrid_p, g.Lx, g.Ly, camera_position, g.arr.cgrid_p,
                          g.Cx, g.Cy, volume_size, volume_position, t0, deltaT, &g.min_idx,
                          &g.max_idx);
    g.pairs = pipeline::precompute_distances(laser_position, g.arr.lgrid_p, g.Lx, g.Ly,
                                             camera_position, g.arr.cgrid_p, g.Cx, g.Cy,
                                             is_confocal != 0);
    return g;
}

// The capture as the device plans take it: the (row-major) source rows with the chunk window
// of get_transient_chunk, cropped on the device (no host chunk); t0 of the chunk is
// chunked_t0 = t0 + min_T_index * deltaT, unfloored (backproject.hpp:847).
CaptureIn device_chunk(const PipelineGeometry& g, float t0, float deltaT) {
    const pipeline::ChunkWindow w = pipeline::chunk_window(g.T, g.min_idx, g.max_idx);
    if (g.rows != g.pairs.size())
        fail(BPX_ERR_INVALID_ARGUMENT, "transient_size (" + std::to_string(g.rows * w.Tc) + ") != num_pairs * T (" +
                                           std::to_string((uint64_t)g.pairs.size() * w.Tc) + ")");
    CaptureIn c{nullptr, g.rows * w.Tc, w.Tc, g.pairs.data(), (uint32_t)g.pairs.size(),
                t0 + g.min_idx * deltaT, deltaT, false};
    c.src = g.arr.data_p;
    c.src_T = g.T;
    c.from = w.from;
    c.cfrom = w.cfrom;
    c.ncopy = w.ncopy;
    return c;
}

}  // namespace

extern "C" {

bpx_status bpx_nlos_backproject(const float* transient_data, const int64_t* data_shape, int data_ndim,
                                const float* camera_grid_positions, const int64_t* camera_grid_shape,
                                const float* laser_grid_positions, const int64_t* laser_grid_shape,
                                const float camera_position[3], const float laser_position[3],
                                float t0, float deltaT, int is_confocal,
                                const float volume_position[3], const float volume_size[3],
                                const uint32_t voxels_per_side[3], const float wall_normal[3],
                                int assume_row_major, int use_cpu, float* volume_out) {
    clear_error();
    try {
        if (!volume_out) fail(BPX_ERR_INVALID_ARGUMENT, "NULL output");
        PipelineGeometry g = prepare(transient_data, data_shape, data_ndim, camera_grid_positions,
                                     camera_grid_shape, laser_grid_positions, laser_grid_shape,
                                     camera_position, laser_position, t0, deltaT, is_confocal,
                                     volume_position, volume_size, voxels_per_side, assume_row_major);
        const CaptureIn c = device_chunk(g, t0, deltaT);
        if (use_cpu) {
            if (!volume_position || !volume_size) fail(BPX_ERR_INVALID_ARGUMENT, "NULL volume argument");
            run_plan_to_host(make_plan_window(c, voxels_per_side, volume_position, volume_size), volume_out);
            return BPX_OK;
        }
        for (int k = 0; k < 3; ++k)
            if (voxels_per_side[k] < 2)
                fail(BPX_ERR_INVALID_ARGUMENT,
                     "the GPU path's grid spacing size/(N-1) needs at least 2 voxels per side");
        const float zeros3[3] = {0, 0, 0};
        float zero[3], inc[9];
        pipeline::basis(volume_position, volume_size, voxels_per_side, wall_normal ? wall_normal : zeros3,
                        zero, inc);
        run_plan_to_host(make_plan_single_tap(c, voxels_per_side, zero, inc, nullptr), volume_out);
        return BPX_OK;
    } catch (const Failure& e) {
        return set_error(e.status, e.msg);
    } catch (const std::bad_alloc&) {
        return set_error(BPX_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
}

bpx_status bpx_nlos_plan_create(const float* transient_data, const int64_t* data_shape, int data_ndim,
                                const float* camera_grid_positions, const int64_t* camera_grid_shape,
                                const float* laser_grid_positions, const int64_t* laser_grid_shape,
                                const float camera_position[3], const float laser_position[3],
                                float t0, float deltaT, int is_confocal,
                                const float volume_position[3], const float volume_size[3],
                                const uint32_t voxels_per_side[3], const float wall_normal[3],
                                int assume_row_major, const bpx_options* options,
                                bpx_plan* plan_out) {
    clear_error();
    try {
        if (!plan_out) fail(BPX_ERR_INVALID_ARGUMENT, "NULL plan_out");
        *plan_out = nullptr;
        PipelineGeometry g = prepare(transient_data, data_shape, data_ndim, camera_grid_positions,
                                     camera_grid_shape, laser_grid_positions, laser_grid_shape,
                                     camera_position, laser_position, t0, deltaT, is_confocal,
                                     volume_position, volume_size, voxels_per_side, assume_row_major);
        const CaptureIn c = device_chunk(g, t0, deltaT);
        for (int k = 0; k < 3; ++k)
            if (voxels_per_side[k] < 2)
                fail(BPX_ERR_INVALID_ARGUMENT,
                     "the GPU path's grid spacing size/(N-1) needs at least 2 voxels per side");
        const float zeros3[3] = {0, 0, 0};
        float zero[3], inc[9];
        pipeline::basis(volume_position, volume_size, voxels_per_side, wall_normal ? wall_normal : zeros3,
                        zero, inc, options && (options->flags & BPX_FLAG_CENTRED_WALL_BASIS));
        auto p = std::make_unique<bpx_plan_s>();
        p->impl.reset(make_plan_single_tap(c, voxels_per_side, zero, inc, options));
        *plan_out = p.release();
        return BPX_OK;
    } catch (const Failure& e) {
        return set_error(e.status, e.msg);
    } catch (const std::bad_alloc&) {
        return set_error(BPX_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
}

bpx_status bpx_nlos_prepare(const float* transient_data, const int64_t* data_shape, int data_ndim,
                            const float* camera_grid_positions, const int64_t* camera_grid_shape,
                            const float* laser_grid_positions, const int64_t* laser_grid_shape,
                            const float camera_position[3], const float laser_position[3],
                            float t0, float deltaT, int is_confocal,
                            const float volume_position[3], const float volume_size[3],
                            const uint32_t voxels_per_side[3], const float wall_normal[3],
                            int assume_row_major, bpx_nlos_prepared* out) {
    clear_error();
    try {
        if (!out) fail(BPX_ERR_INVALID_ARGUMENT, "NULL output");
        std::memset(out, 0, sizeof(*out));
        PipelineGeometry g = prepare(transient_data, data_shape, data_ndim, camera_grid_positions,
                                     camera_grid_shape, laser_grid_positions, laser_grid_shape,
                                     camera_position, laser_position, t0, deltaT, is_confocal,
                                     volume_position, volume_size, voxels_per_side, assume_row_major);
        uint32_t Tc = 0;
        std::vector<float> chunk =
            pipeline::transient_chunk(g.arr.data_p, g.rows, g.T, g.min_idx, g.max_idx, &Tc);
        const float zeros3[3] = {0, 0, 0};
        pipeline::basis(volume_position, volume_size, voxels_per_side, wall_normal ? wall_normal : zeros3,
                        out->zero, out->inc);
        out->chunk = static_cast<float*>(std::malloc(sizeof(float) * std::max<size_t>(chunk.size(), 1)));
        out->pairs = static_cast<bpx_pair*>(std::malloc(sizeof(bpx_pair) * std::max<size_t>(g.pairs.size(), 1)));
        if (!out->chunk || !out->pairs) {
            bpx_nlos_prepared_free(out);
            fail(BPX_ERR_OUT_OF_MEMORY, "host allocation failed");
        }
        std::memcpy(out->chunk, chunk.data(), sizeof(float) * chunk.size());
        std::memcpy(out->pairs, g.pairs.data(), sizeof(bpx_pair) * g.pairs.size());
        out->rows = g.rows;
        out->Tc = Tc;
        out->num_pairs = (uint32_t)g.pairs.size();
        out->chunked_t0 = t0 + g.min_idx * deltaT;
        out->min_index = g.min_idx;
        out->max_index = g.max_idx;
        return BPX_OK;
    } catch (const Failure& e) {
        return set_error(e.status, e.msg);
    } catch (const std::bad_alloc&) {
        return set_error(BPX_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
}

void bpx_nlos_prepared_free(bpx_nlos_prepared* p) {
    if (!p) return;
    std::free(p->chunk);
    std::free(p->pairs);
    p->chunk = nullptr;
    p->pairs = nullptr;
}

bpx_status bpx_nlos_phasor(const float* transient_data, const int64_t* data_shape, int data_ndim,
                           const float* camera_grid_positions, const int64_t* camera_grid_shape,
                           const float* laser_grid_positions, const int64_t* laser_grid_shape,
                           const float camera_position[3], const float laser_position[3],
                           float t0, float deltaT, int is_confocal,
                           const float volume_position[3], const float volume_size[3],
                           const uint32_t voxels_per_side[3], float wavelength,
                           int assume_row_major, int use_cpu, float* complex_volume_out) {
    return bpx_nlos_phasor_multi(transient_data, data_shape, data_ndim, camera_grid_positions, camera_grid_shape,
                                 laser_grid_positions, laser_grid_shape, camera_position, laser_position, t0,
                                 deltaT, is_confocal, volume_position, volume_size, voxels_per_side, &wavelength, 1u,
                                 assume_row_major, use_cpu, complex_volume_out);
}

bpx_status bpx_nlos_phasor_multi(const float* transient_data, const int64_t* data_shape, int data_ndim,
                                 const float* camera_grid_positions, const int64_t* camera_grid_shape,
                                 const float* laser_grid_positions, const int64_t* laser_grid_shape,
                                 const float camera_position[3], const float laser_position[3],
                                 float t0, float deltaT, int is_confocal,
                                 const float volume_position[3], const float volume_size[3],
                                 const uint32_t voxels_per_side[3], const float* wavelengths, uint32_t count,
                                 int assume_row_major, int use_cpu, float* complex_volumes_out) {
    return bpx_nlos_phasor_output(transient_data, data_shape, data_ndim, camera_grid_positions, camera_grid_shape,
                                  laser_grid_positions, laser_grid_shape, camera_position, laser_position, t0,
                                  deltaT, is_confocal, volume_position, volume_size, voxels_per_side, wavelengths,
                                  count, assume_row_major, use_cpu, BPX_PHASOR_COMPLEX, complex_volumes_out);
}

bpx_status bpx_nlos_phasor_output(const float* transient_data, const int64_t* data_shape, int data_ndim,
                                  const float* camera_grid_positions, const int64_t* camera_grid_shape,
                                  const float* laser_grid_positions, const int64_t* laser_grid_shape,
                                  const float camera_position[3], const float laser_position[3],
                                  float t0, float deltaT, int is_confocal,
                                  const float volume_position[3], const float volume_size[3],
                                  const uint32_t voxels_per_side[3], const float* wavelengths, uint32_t count,
                                  int assume_row_major, int use_cpu, int output, float* volumes_out) {
    clear_error();
    try {
        if (!volumes_out) fail(BPX_ERR_INVALID_ARGUMENT, "NULL output");
        if (output < BPX_PHASOR_COMPLEX || output > BPX_PHASOR_MAGNITUDE_PHASE)
            fail(BPX_ERR_INVALID_ARGUMENT, "output must be a bpx_phasor_output value");
        if (count == 0) return BPX_OK;
        if (!wavelengths) fail(BPX_ERR_INVALID_ARGUMENT, "NULL wavelengths");
        PipelineGeometry g = prepare(transient_data, data_shape, data_ndim, camera_grid_positions,
                                     camera_grid_shape, laser_grid_positions, laser_grid_shape,
                                     camera_position, laser_position, t0, deltaT, is_confocal,
                                     volume_position, volume_size, voxels_per_side, assume_row_major);
        // :962-969 -- the float limits are truncated to int
        const float imin = (float)(int)g.min_idx, imax = (float)(int)g.max_idx;
        const pipeline::ChunkWindow w = pipeline::chunk_window(g.T, imin, imax);
        std::vector<std::vector<float>> cs(count), sn(count);
        for (uint32_t i = 0; i < count; ++i)
            if (pipeline::phasor_taps(wavelengths[i], deltaT, 4, 0.0f, cs[i], sn[i]) <= 0)
                fail(BPX_ERR_INVALID_ARGUMENT, "the phasor pulse has no taps for wavelength " +
                                                   std::to_string(wavelengths[i]));
        if (!device_ready()) fail(BPX_ERR_NO_DEVICE, "no gfx950 device available");
        if (g.rows != g.pairs.size())
            fail(BPX_ERR_INVALID_ARGUMENT, "transient_size (" + std::to_string(g.rows * w.Tc) +
                                               ") != num_pairs * T (" +
                                               std::to_string((uint64_t)g.pairs.size() * w.Tc) + ")");
        if (!use_cpu)
            for (int k = 0; k < 3; ++k)
                if (voxels_per_side[k] < 2)
                    fail(BPX_ERR_INVALID_ARGUMENT,
                         "the GPU path's grid spacing size/(N-1) needs at least 2 voxels per side");
        // one upload and crop; per wavelength the FIR, then the same plan re-staged and run
        const DeviceBuffer chunk = chunk_rows_device(g.arr.data_p, g.rows, g.T, w);
        const float chunked_t0 = t0 + imin * deltaT;  // :1011
        const size_t nvox = (size_t)voxels_per_side[0] * voxels_per_side[1] * voxels_per_side[2];
        std::unique_ptr<PlanImpl> plan;
        // floats per voxel and wavelength in the output, and the device buffer of the real outputs
        const size_t per = output == BPX_PHASOR_COMPLEX || output == BPX_PHASOR_MAGNITUDE_PHASE ? 2 : 1;
        DeviceBuffer post;
        if (output != BPX_PHASOR_COMPLEX) post.alloc(4 * nvox * per);
        for (uint32_t i = 0; i < count; ++i) {
            DeviceBuffer rows = phasor_rows_device(chunk, g.rows, w.Tc, cs[i], sn[i]);
            if (!plan) {
                CaptureIn c{nullptr, g.rows * w.Tc, w.Tc, g.pairs.data(), (uint32_t)g.pairs.size(), chunked_t0,
                            deltaT, true};
                c.device_rows = &rows;
                if (use_cpu) {
                    if (!volume_position || !volume_size) fail(BPX_ERR_INVALID_ARGUMENT, "NULL volume argument");
                    plan.reset(make_plan_window(c, voxels_per_side, volume_position, volume_size));
                } else {
                    const float zeros3[3] = {0, 0, 0};
                    float zero[3], inc[9];
                    pipeline::basis(volume_position, volume_size, voxels_per_side, zeros3, zero, inc);
                    plan.reset(make_plan_single_tap(c, voxels_per_side, zero, inc, nullptr));
                }
            } else {
                plan_replace_rows(*plan, rows);
            }
            plan->execute(plan->stream);
            const void* src = plan->vol.ptr;
            if (output != BPX_PHASOR_COMPLEX) {
                float* mag = static_cast<float*>(post.ptr);
                float* ph = output == BPX_PHASOR_MAGNITUDE_PHASE ? mag + nvox : mag;
                hip_check(launch_complex_post(plan->vol.ptr, nvox, output == BPX_PHASOR_PHASE ? nullptr : mag,
                                              output == BPX_PHASOR_MAGNITUDE ? nullptr : ph, plan->stream),
                          "complex post-processing kernel");
                src = post.ptr;
            }
            hip_check(hipMemcpyAsync(volumes_out + per * nvox * i, src, 4 * per * nvox, hipMemcpyDeviceToHost,
                                     plan->stream),
                      "hipMemcpyAsync(volume)");
            hip_check(hipStreamSynchronize(plan->stream), "hipStreamSynchronize");
        }
        return BPX_OK;
    } catch (const Failure& e) {
        return set_error(e.status, e.msg);
    } catch (const std::bad_alloc&) {
        return set_error(BPX_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
}

}  // extern "C"
