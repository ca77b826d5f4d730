// Synthetic transient renderer (no reference counterpart): hidden Lambertian samples seen
// through the relay wall, binned by total path length, optional Poisson photon noise.
// Deterministic: each pair's row is accumulated sequentially over the samples, and every
// Poisson draw comes from a counter-based generator keyed by (seed, pair, bin).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>
#include <vector>

#include "bpx_host.hpp"

namespace bpx {

static inline uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

struct Counter {
    uint64_t key, n = 0;
    double uniform() {  // (0, 1)
        const uint64_t r = splitmix64(key ^ splitmix64(++n));
        return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    }
};

// Poisson draw: inversion below 10, Hormann's transformed rejection (PTRS) above.
static double poisson(double lam, Counter& rng) {
    if (!(lam > 0.0)) return 0.0;
    if (lam < 10.0) {
        double p = std::exp(-lam), s = p, u = rng.uniform();
        double k = 0.0;
        while (u > s && k < 1000.0) {
            k += 1.0;
            p *= lam / k;
            s += p;
        }
        return k;
    }
    const double slam = std::sqrt(lam), loglam = std::log(lam);
    const double b = 0.931 + 2.53 * slam, a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4), vr = 0.9277 - 3.6224 / (b - 2.0);
    for (;;) {
        const double U = rng.uniform() - 0.5, V = rng.uniform();
        const double us = 0.5 - std::fabs(U);
        const double k = std::floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return k;
        if (k < 0.0 || (us < 0.013 && V > us)) continue;
        if (std::log(V) + std::log(invalpha) - std::log(a / (us * us) + b) <=
            -lam + k * loglam - std::lgamma(k + 1.0))
            return k;
    }
}

template <class F>
static void parallel_for(uint64_t n, F&& f) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 64) nt = 1;
    std::atomic<uint64_t> next{0};
    std::vector<std::thread> th;
    auto work = [&] {
        for (;;) {
            const uint64_t i = next.fetch_add(1);
            if (i >= n) break;
            f(i);
        }
    };
    for (unsigned t = 1; t < nt; ++t) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
}

void render_transient(const bpx_pair* pairs, uint32_t P, uint32_t T, float t0, float dT,
                      const float* samples, uint64_t n, float peak, bool use_poisson, uint64_t seed,
                      float* out) {
    std::vector<double> rowmax(P, 0.0);
    parallel_for(P, [&](uint64_t p) {
        std::vector<double> row(T, 0.0);
        const bpx_pair& pr = pairs[p];
        for (uint64_t i = 0; i < n; ++i) {
            const float* q = samples + 7 * i;
            double vl[3], vs[3];
            for (int k = 0; k < 3; ++k) {
                vl[k] = (double)pr.laser_point[k] - q[k];
                vs[k] = (double)pr.camera_point[k] - q[k];
            }
            const double dl2 = vl[0] * vl[0] + vl[1] * vl[1] + vl[2] * vl[2];
            const double ds2 = vs[0] * vs[0] + vs[1] * vs[1] + vs[2] * vs[2];
            if (!(dl2 > 0.0) || !(ds2 > 0.0)) continue;
            const double dl = std::sqrt(dl2), ds = std::sqrt(ds2);
            const double L = (double)pr.laser_wall + dl + ds + (double)pr.camera_wall;
            const double fb = std::floor((L - (double)t0) / (double)dT);
            if (!(fb >= 0.0) || fb >= (double)T) continue;
            double w = (double)q[6] / (dl2 * ds2);
            const double nn = (double)q[3] * q[3] + (double)q[4] * q[4] + (double)q[5] * q[5];
            if (nn > 0.0) {
                const double cl = (q[3] * vl[0] + q[4] * vl[1] + q[5] * vl[2]) / (dl * std::sqrt(nn));
                const double cs = (q[3] * vs[0] + q[4] * vs[1] + q[5] * vs[2]) / (ds * std::sqrt(nn));
                if (cl <= 0.0 || cs <= 0.0) continue;
                w *= cl * cs;
            }
            row[(size_t)fb] += w;
        }
        double m = 0.0;
        for (uint32_t b = 0; b < T; ++b) {
            out[p * T + b] = (float)row[b];
            m = std::max(m, row[b]);
        }
        rowmax[p] = m;
    });
    double m = 0.0;
    for (double v : rowmax) m = std::max(m, v);
    const double scale = m > 0.0 ? (double)peak / m : 0.0;
    parallel_for(P, [&](uint64_t p) {
        for (uint32_t b = 0; b < T; ++b) {
            const double lam = (double)out[p * T + b] * scale;
            if (use_poisson) {
                Counter rng{splitmix64(seed ^ splitmix64(p * 0x100000001b3ull + b))};
                out[p * T + b] = (float)poisson(lam, rng);
            } else {
                out[p * T + b] = (float)lam;
            }
        }
    });
}

}  // namespace bpx
