// icm_common.hpp -- host utilities shared by the plan and the ragged-batch
// plan: error state, device selection, quadrature nodes, tree geometry and
// the per-launch event profiler.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/icm_hip.h"
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

// Geometry of the small-instance kernel (blocks of one player); returns the
// LDS doubles it needs (or a value above every tier when it cannot run).
inline int small_geometry(int n, int k, icm::SmallJob &J) {
    J.n = n;
    J.k = k;
    int L = 1;
    while ((1LL << (L - 1)) < n) ++L;
    J.L = L;
    if (L > icm::kSmallMaxL) return 1 << 30;
    long long off = n, gmax = 0;
    J.nr[0] = n;
    for (int l = 0; l < L; ++l) {
        if (l > 0) J.nr[l] = (J.nr[l - 1] + 1) / 2;
        J.c[l] = (int)std::min<long long>((1LL << l) + 1, k);
        J.g[l] = (int)std::min<long long>(1LL << l, k);
        J.off[l] = 0;
        if (l >= 1 && l <= L - 2) {
            J.off[l] = (int)off;
            off += (long long)J.nr[l] * J.c[l];
        }
        if (l <= L - 2) gmax = std::max(gmax, (long long)J.nr[l] * J.g[l]);
    }
    J.goff[0] = (int)off;
    J.goff[1] = (int)(off + gmax);
    const long long need = off + 2 * gmax;
    J.lds_doubles = (int)std::min<long long>(need, 1 << 30);
    return J.lds_doubles;
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
    struct Rec { std::string name; hipEvent_t a, b; double bytes, flops; int tag; };
    struct Agg { std::string name; int launches = 0; double ns = 0, bytes = 0, flops = 0; };
    struct Launch { std::string name; int tag; double ns, bytes, flops; };
    bool enabled = false;
    int tag = 0;                        // set by the caller before launches: tree level (+ build, - propagate)
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    std::vector<Rec> pending;
    std::vector<Agg> agg;
    std::vector<Launch> log;            // every launch since profiling was switched on
    hipEvent_t cur_a = nullptr;
    void clear() {
        agg.clear();
        pending.clear();
        log.clear();
        used = 0;
    }
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
        pending.push_back(Rec{name, cur_a, b, bytes, flops, tag});
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
            if (log.size() < (1u << 20)) log.push_back(Launch{r.name, r.tag, (double)ms * 1e6, r.bytes, r.flops});
        }
        pending.clear();
        used = 0;
    }
};


// Plan configuration from the reference options struct, the engine's own
// parameters and the environment (ICM_GPU_Q_BATCH as src/gpu/gpu_api.cu:139-144;
// ICM_HIP_MEM_BUDGET, bytes; ICM_HIP_MEMORY_STRATEGY overriding
// memory_strategy, as bench/bench_gpu.cu:60 does for its options).  Values
// this engine cannot honour are errors.
struct PlanCfg {
    int B = 0;               // forced block size (0 = tuning table)
    int qb = 0;              // forced q-planes per batch (0 = from the budget)
    double frac = 0.6;       // share of free device memory to plan for
    size_t cap = 0;          // device-byte cap (0 = none)
    bool spec_ok = true;     // spectrum cache allowed (memory_strategy != 3)
    bool verbose = false;
    bool graphs = true;      // ragged batches: record / replay a HIP graph
    int ckpt = -1;           // adjoint checkpoint interval (-1 auto, 0 keep every level, 2 .. 64)
    bool batch_persistent = false;  // ragged batches: one persistent launch per q-pass
};

inline bool plan_config(const IcmHipOptions *o, const IcmHipPlanParams *pp, PlanCfg &c) {
    char buf[192];
    if (o) {
        if (o->force_uncached_fused_levels > 0 || o->force_uncached_cufft_levels > 0) {
            snprintf(buf, sizeof buf,
                     "force_uncached_fused_levels=%d / force_uncached_cufft_levels=%d: this engine has no "
                     "per-level cache count (use -1); block size and q-batch are IcmHipPlanParams fields",
                     o->force_uncached_fused_levels, o->force_uncached_cufft_levels);
            set_err(buf);
            return false;
        }
        c.spec_ok = o->memory_strategy != 3;
        c.verbose = o->verbose != 0;
        c.graphs = o->enable_graphs != 0;
    }
    if (pp) {
        if (pp->block_size != 0 && pp->block_size != 16 && pp->block_size != 32 && pp->block_size != 64) {
            snprintf(buf, sizeof buf, "block_size=%d: expected 0 (tuned), 16, 32 or 64", pp->block_size);
            set_err(buf);
            return false;
        }
        if (pp->q_batch < 0 || pp->q_batch > 256) {
            snprintf(buf, sizeof buf, "q_batch=%d: expected 0 (auto) .. 256", pp->q_batch);
            set_err(buf);
            return false;
        }
        if (!(pp->mem_fraction >= 0.0 && pp->mem_fraction <= 1.0)) {
            set_err("mem_fraction must lie in [0, 1] (0 = default 0.6)");
            return false;
        }
        c.B = pp->block_size;
        c.qb = pp->q_batch;
        if (pp->mem_fraction > 0.0) c.frac = pp->mem_fraction;
        c.cap = pp->mem_budget_bytes;
    }
    if (c.qb == 0) {
        const char *e = getenv("ICM_GPU_Q_BATCH");
        if (e && e[0]) {
            const int v = atoi(e);
            if (v >= 1 && v <= 256) c.qb = v;
        }
    }
    {   // memory strategy override, as the reference bench's ICM_GPU_BENCH_MEMORY_STRATEGY
        const char *e = getenv("ICM_HIP_MEMORY_STRATEGY");
        if (e && e[0]) c.spec_ok = atoi(e) != 3;
    }
    if (c.cap == 0) {
        const char *e = getenv("ICM_HIP_MEM_BUDGET");
        if (e && e[0]) c.cap = (size_t)strtoull(e, nullptr, 10);
    }
    {
        const char *e = getenv("ICM_HIP_BATCH_PERSISTENT");
        if (e && e[0]) c.batch_persistent = atoi(e) != 0;
    }
    {
        const char *e = getenv("ICM_HIP_ADJOINT_CHECKPOINT");
        if (e && e[0]) {
            const int v = atoi(e);
            if (v == -1 || v == 0 || (v >= 2 && v <= 64)) c.ckpt = v;
        }
    }
    return true;
}

// Device allocation accounting shared by the plan types: a byte counter, an
// optional cap (an allocation beyond it fails as an out-of-memory would) and
// a dry mode that only counts (host-only planning).  A failed hipMalloc's
// error is cleared so that later launches do not report it.
struct DevMem {
    size_t bytes = 0;
    size_t cap = 0;
    bool dry = false;
    template <class T>
    bool alloc(T **ptr, size_t count) {
        if (count == 0) count = 1;
        const size_t b = count * sizeof(T);
        *ptr = nullptr;
        if (cap && bytes + b > cap) {
            char buf[160];
            snprintf(buf, sizeof buf, "device memory budget exceeded: %.3f GB held, %.3f GB more, cap %.3f GB",
                     bytes / 1e9, b / 1e9, cap / 1e9);
            set_err(buf);
            return false;
        }
        if (!dry) {
            const hipError_t e = hipMalloc((void **)ptr, b);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                *ptr = nullptr;
                hip_ok(e, "hipMalloc");
                return false;
            }
        }
        bytes += b;
        return true;
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
