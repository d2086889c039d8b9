// icm_common.hpp -- host utilities shared by the plan and the ragged-batch
// plan: error state, device selection, quadrature nodes, tree geometry and
// the per-launch event profiler.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "icm_kernels.hpp"

namespace icmh {

inline thread_local std::string g_err;
inline int g_device = -1;
inline std::mutex g_mu;

inline void set_err(const std::string &s) { g_err = s; }

inline bool hip_ok(hipError_t e, const char *what) {
    if (e == hipSuccess) return true;
    char buf[256];
    snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    set_err(buf);
    return false;
}
#define HIP_OK(x) hip_ok((x), #x)

inline int next_pow2_ll(long long v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}
inline int ilog2i(long long v) {
    int l = 0;
    while ((1LL << l) < v) ++l;
    return l;
}

// erfc-trap nodes: src/cpu/icm.c:117-132 (h = 0 for Q = 1 as src/gpu/gpu_exec.cu:17).
inline double log_Phi(double y) {
    if (y >= 0) return std::log1p(-std::erfc(y / std::sqrt(2.0)) / 2.0);
    return std::log(std::erfc(-y / std::sqrt(2.0)) / 2.0);
}
inline void make_nodes(int Q, double Smax, double *logv, double *w) {
    const double y_lo = -7.7;
    double y_hi = std::sqrt(2.0) * std::sqrt(std::log(Smax) + 25.0);
    if (y_hi < 6.5) y_hi = 6.5;
    const double h = (Q > 1) ? (y_hi - y_lo) / (Q - 1) : 0.0;
    for (int q = 0; q < Q; ++q) {
        const double y = y_lo + q * h;
        logv[q] = log_Phi(y);
        w[q] = h * std::exp(-0.5 * y * y) / std::sqrt(2.0 * M_PI);
    }
}

struct Geometry {
    int n = 0, k = 0, B = 0, nb = 0, L = 0;
    std::vector<int> n_real, c, G, F;
};

inline bool make_geometry(int n, int k, int B, Geometry &g) {
    if (n < 1 || k < 1 || k > n || (B != 16 && B != 32 && B != 64)) return false;
    g.n = n; g.k = k; g.B = B;
    g.nb = (n + B - 1) / B;
    const int N = next_pow2_ll(g.nb);
    g.L = ilog2i(N) + 1;
    g.n_real.assign(g.L, 0); g.c.assign(g.L, 0); g.G.assign(g.L, 0); g.F.assign(g.L, 0);
    g.n_real[0] = g.nb;
    for (int l = 1; l < g.L; ++l) g.n_real[l] = (g.n_real[l - 1] + 1) / 2;
    for (int l = 0; l < g.L; ++l) {
        const long long D = (long long)B << l;
        g.c[l] = (int)std::min<long long>(D + 1, k);
        g.G[l] = (int)std::min<long long>(D, k);
    }
    for (int l = 1; l < g.L; ++l) {
        const long long need = std::max<long long>(
            {2LL * g.c[l - 1] - 2, (long long)g.G[l - 1] + g.c[l - 1] - 1, 16LL});
        g.F[l] = next_pow2_ll(need);
    }
    return true;
}

// Schoolbook / transform crossover: levels whose children keep at most this
// many coefficients run as direct sums.  Chosen by icm_amd/autotune.py on
// MI355X (tuning table in tuning_mi355x.h).
#include "tuning_mi355x.h"

inline const TuneRow &tune_row(int n, int k) {
    for (const TuneRow &r : kTuneTable)
        if (n <= r.n_max && k <= r.k_max) return r;
    return kTuneTable[sizeof(kTuneTable) / sizeof(kTuneTable[0]) - 1];
}

inline int tuned_schoolbook_max(int n, int k) { return tune_row(n, k).sb_max; }

inline int pick_B(int n, int k, int forced) {
    if (forced == 16 || forced == 32 || forced == 64) return forced;
    const int B = tune_row(n, k).B;
    return (B == 16 || B == 32 || B == 64) ? B : 32;
}

// Per-launch event timing, aggregated per kernel name (see icm::Prof).
struct KernelProf : icm::Prof {
    struct Rec { std::string name; hipEvent_t a, b; double bytes, flops; };
    struct Agg { std::string name; int launches = 0; double ns = 0, bytes = 0, flops = 0; };
    bool enabled = false;
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    std::vector<Rec> pending;
    std::vector<Agg> agg;
    hipEvent_t cur_a = nullptr;
    ~KernelProf() override { for (hipEvent_t e : pool) (void)hipEventDestroy(e); }
    bool on() const override { return enabled; }
    hipEvent_t take() {
        if (used == pool.size()) {
            hipEvent_t e = nullptr;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            pool.push_back(e);
        }
        return pool[used++];
    }
    void begin(hipStream_t s) override {
        cur_a = take();
        if (cur_a) (void)hipEventRecord(cur_a, s);
    }
    void end(hipStream_t s, const char *name, double bytes, double flops) override {
        hipEvent_t b = take();
        if (!cur_a || !b) return;
        (void)hipEventRecord(b, s);
        pending.push_back(Rec{name, cur_a, b, bytes, flops});
        cur_a = nullptr;
    }
    // After the stream has finished: fold pending records into the aggregates.
    void collect() {
        for (const Rec &r : pending) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) continue;
            Agg *g = nullptr;
            for (Agg &x : agg) if (x.name == r.name) { g = &x; break; }
            if (!g) { agg.push_back(Agg{}); g = &agg.back(); g->name = r.name; }
            g->launches += 1;
            g->ns += (double)ms * 1e6;
            g->bytes += r.bytes;
            g->flops += r.flops;
        }
        pending.clear();
        used = 0;
    }
};


inline bool ensure_device() {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_device >= 0) return HIP_OK(hipSetDevice(g_device));
    int cnt = 0;
    if (!HIP_OK(hipGetDeviceCount(&cnt))) return false;
    if (cnt < 1) { set_err("no HIP device"); return false; }
    if (!HIP_OK(hipSetDevice(0))) return false;
    g_device = 0;
    return true;
}

}  // namespace icmh
