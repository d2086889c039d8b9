// icm_host.cpp -- planner, plan lifecycle and the C ABI of libicm_hip.so.
//
// Tree geometry (this engine's own; the reference's is src/cpu/icm.c:1205-1263
// and src/gpu/gpu_cost_model.cu:494-572):
//   nb = ceil(n / B) blocks, N = next_pow2(nb), L = log2(N) + 1 levels,
//   n_real[0] = nb, n_real[l] = ceil(n_real[l-1] / 2)   (ragged: no padding nodes)
//   D_l = B 2^l   (largest degree of a level-l node)
//   c_l = min(D_l + 1, k)   coefficients kept per node (product mod x^k)
//   G_l = min(D_l, k)       g entries per node (pairs with leave-one-out
//                            products of degree <= D_l - 1, truncated at k)
//   F_l = next_pow2(max(2 c_{l-1} - 2, G_{l-1} + c_{l-1} - 1, 16))
// F_l serves both passes of level l: the build product of two c_{l-1}-term
// polynomials aliases at most coefficient 0 (fixed exactly from the children),
// and the propagate correlation g_child[m] = sum_t P_sib[t] g_par[m + t],
// m < G_{l-1}, never wraps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/icm_hip.h"
#include "icm_common.hpp"
#include "icm_kernels.hpp"

using namespace icmh;


struct IcmHipPlan {
    Geometry geo;
    int qb = 0;                         // q-planes per batch
    int nchunk = 1;                     // q-chunks of the leaf extract
    int twShift = 0;
    int sb_max = 0;                     // levels with c_{l-1} <= sb_max use schoolbook sums
    std::vector<int> perm;              // sorted position -> original index
    std::vector<double> S_sorted;
    double Smax = 0.0;
    bool verbose = false;
    // device memory
    double2 *d_tw = nullptr;
    double *d_S = nullptr;
    int *d_perm = nullptr;
    double *d_payout = nullptr;          // k (copied per call)
    double *d_nodes = nullptr;           // [2][Qcap]: logv, w of the active nodes
    int q_cap = 0;
    int nodes_Q = 0;                     // Q the uploaded nodes belong to (-1: the raw per-node entry)
    int nodes_Qa = 0;                    // active nodes (w != 0) among them
    // small instances: the whole quadrature in one launch (k_small), chosen
    // when the tree fits one workgroup's LDS and the tuned rule allows it
    int small_mode = -1;                 // -1 auto, 0 off, 1 whenever it fits
    bool small_fits = false;
    bool small_auto = false;             // the tuned rule's choice
    icm::SmallJob sj{};
    int small_gmax = 0;                  // workgroups (rows of d_small)
    double *d_small = nullptr;           // [small_gmax][n] per-workgroup partial equities
    double *d_blocks = nullptr;          // [qb][nb][B+1]
    std::vector<double *> d_coef;        // level l (1..L-2): [qb][n_real][c]
    double *d_g[2] = {nullptr, nullptr};
    long long g_elems = 0;               // per q-plane
    double2 *d_bufA = nullptr, *d_bufB = nullptr;
    long long big_elems = 0;             // complex per q-plane per buffer
    // spectrum cache: column transforms of the child pairs of every four-step
    // build level, kept from the build for the propagate of the same q-batch
    // (level l at d_spec + spec_off[l], [qb][n_real[l]][F[l]] complex)
    double2 *d_spec = nullptr;
    std::vector<long long> spec_off;
    long long spec_elems = 0;            // complex per q-plane
    size_t spec_bytes = 0;
    double *d_partial = nullptr;         // [nchunk][n]
    double *d_acc = nullptr;             // [n] sorted order
    double *d_equity = nullptr;          // [n] original order
    unsigned char *d_active = nullptr;   // [n] sorted order or null
    // subset pruning (q-independent): hot parents per propagate level and hot
    // blocks, concatenated in d_lists; offsets/counts per level (index L = blocks)
    int *d_lists = nullptr;
    std::vector<int> list_off, list_cnt;
    const double *d_a_in = nullptr;      // optional a values [qb][n]
    bool raw_node = false;               // single node with logv = 0, w = 1 (per-node entry)
    double *d_inner = nullptr;           // optional Psi output [qb][n]
    // adjoint (VJP) buffers, allocated on the first VJP call
    int qv = 0;                          // q-planes per VJP batch
    double *d_dblocks = nullptr;         // [qv][nb][B+1] tangent block products
    std::vector<double *> d_dcoef;       // level l: [qv][n_real][c] tangent products (checkpoint levels only
                                         //   when checkpointing)
    int ckpt_mode = -1;                  // adjoint checkpointing: -1 auto, 0 keep every tangent level,
                                         //   m >= 2 keep levels 0, m, 2m, ... and recompute the rest
    int ckpt = 0;                        // interval in use (0: every tangent level kept)
    std::vector<double *> d_dseg;        // m - 1 slots: the recomputed tangent levels of one segment
    long long seg_elems = 0;             // per q-plane per slot
    size_t vjp_base = 0;                 // device bytes held before the adjoint buffers
    double *d_bv[2] = {nullptr, nullptr};// B adjoint ping-pong, [qv][n_real][c]
    double2 *d_dspec = nullptr;          // tangent spectrum cache: the dP pairs' row-stage inputs,
    std::vector<long long> dspec_off;    //   per four-step build level [qv][n_real[l]][F[l]]
    long long bv_elems = 0;
    double *d_v = nullptr;               // cotangent, sorted order
    double *d_partial_ds = nullptr;      // [nchunk][n]
    double *d_acc_ds = nullptr;          // [n] sorted
    double *d_ds = nullptr;              // [n] original order
    const double *d_c_in = nullptr;      // optional tangent weights (per-node entry)
    double *d_h = nullptr;               // optional H output (per-node entry)
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DevMem mem;                          // device bytes held (cap / dry-run accounting)
    int retries = 0;                     // q-batch halvings at creation (out-of-memory policy)
    int maxF = 16;
    bool spec_ok = true;                 // spectrum cache allowed (memory_strategy != 3)
    double frac = 0.6;                   // share of free device memory to plan for
    int fuse_columns = 15;               // column fusion bits: 1 propagate, 2 build, 4 build into the root,
                                         // 8 the F = 8192 build into the level above
    bool poison = false;                 // diagnostic: NaN-fill the per-batch buffers before every q-batch
    long long poisoned_bytes = 0;        // bytes NaN-filled by the last run
    KernelProf prof;
};

namespace {

template <class T>
bool dalloc(IcmHipPlan *p, T **ptr, size_t count) {
    return p->mem.alloc(ptr, count);
}

// Frees every device buffer of the plan (forward, adjoint, subset lists) and
// clears the byte count; the stream and events stay.
void free_buffers(IcmHipPlan *p) {
    auto f = [](auto *&x) { if (x) (void)hipFree((void *)x); x = nullptr; };
    f(p->d_tw); f(p->d_S); f(p->d_perm); f(p->d_payout); f(p->d_nodes); f(p->d_blocks);
    for (double *&c : p->d_coef) f(c);
    f(p->d_g[0]); f(p->d_g[1]); f(p->d_bufA); f(p->d_bufB); f(p->d_spec); f(p->d_partial); f(p->d_acc);
    f(p->d_equity); f(p->d_active); f(p->d_inner); f(p->d_lists);
    f(p->d_dblocks); for (double *&c : p->d_dcoef) f(c);
    for (double *&c : p->d_dseg) f(c);
    f(p->d_bv[0]); f(p->d_bv[1]); f(p->d_dspec); f(p->d_v); f(p->d_partial_ds); f(p->d_acc_ds); f(p->d_ds);
    f(p->d_small);
    p->ckpt = 0;
    p->q_cap = 0;
    p->nodes_Q = 0;
    p->qv = 0;
    p->spec_bytes = 0;
    p->mem.bytes = 0;
}

void plan_free(IcmHipPlan *p) {
    if (!p) return;
    free_buffers(p);
    if (p->ev0) (void)hipEventDestroy(p->ev0);
    if (p->ev1) (void)hipEventDestroy(p->ev1);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}


// Per-q-plane device bytes of the tree for this geometry.
long long plane_bytes(const Geometry &g, long long *g_elems, long long *big_elems) {
    long long bytes = (long long)g.nb * (g.B + 1) * 8;
    long long ge = 0, be = 0;
    for (int l = 1; l <= g.L - 2; ++l) bytes += (long long)g.n_real[l] * g.c[l] * 8;
    for (int l = 0; l < g.L; ++l) ge = std::max(ge, (long long)g.n_real[l] * g.G[l]);
    for (int l = 1; l < g.L; ++l) {
        if (g.F[l] > icm::kMaxLdsF && !icm::col_fused(g.F[l])) {
            const long long nrp = g.n_real[l], np = (nrp + 1) / 2;
            be = std::max(be, (nrp + np) * (long long)g.F[l]);
        }
    }
    bytes += 2 * ge * 8 + 2 * be * 16;
    *g_elems = ge;
    *big_elems = be;
    return bytes;
}

// Four-step build levels with column arrays (the spectrum cache keeps them).
static bool spec_level(const Geometry &g, int l) {
    return l >= 1 && l <= g.L - 2 && g.F[l] > icm::kMaxLdsF && !icm::col_fused(g.F[l]);
}

// Complex elements per q-plane of the spectrum cache.
long long spec_plane_elems(const Geometry &g) {
    long long e = 0;
    for (int l = 1; l <= g.L - 2; ++l)
        if (spec_level(g, l)) e += (long long)g.n_real[l] * g.F[l];
    return e;
}

bool spec_alloc(IcmHipPlan *p) {
    const Geometry &g = p->geo;
    p->spec_off.assign(g.L, -1);
    long long off = 0;
    for (int l = 1; l <= g.L - 2; ++l) {
        if (spec_level(g, l)) {
            p->spec_off[l] = off;
            off += (long long)p->qb * g.n_real[l] * g.F[l];
        }
    }
    if (off == 0) return true;
    if (!dalloc(p, &p->d_spec, (size_t)off)) return false;
    p->spec_bytes = sizeof(double2) * (size_t)off;
    return true;
}

void spec_release(IcmHipPlan *p) {
    if (!p->d_spec) return;
    (void)hipFree(p->d_spec);
    p->d_spec = nullptr;
    p->mem.bytes -= p->spec_bytes;
    p->spec_bytes = 0;
}

// Schoolbook multiply-adds per quadrature node of the small-instance kernel.
double small_work(const icm::SmallJob &J) {
    double w = 0.0;
    for (int l = 1; l <= J.L - 2; ++l) w += (double)J.nr[l] * J.c[l] * std::min(J.c[l - 1], J.c[l]);
    for (int l = J.L - 1; l >= 1; --l) w += (double)J.nr[l - 1] * J.g[l - 1] * J.c[l - 1];
    return w;
}

// Allocates the forward buffers for qb q-planes (spectrum cache if spec).
bool alloc_buffers(IcmHipPlan *p, int qb, bool spec) {
    const Geometry &g = p->geo;
    const int n = g.n, k = g.k, B = g.B;
    p->qb = qb;
    // leaf-extract q-chunks: enough workgroups to fill the chip several times
    // over (about 16k; 4k left a part-filled last round at n = 2^20)
    const int bpw = 256 / B;
    const int wg_blocks = (g.nb + bpw - 1) / bpw;
    const int nchunk = std::max(1, 16384 / std::max(1, wg_blocks));
    p->nchunk = std::min(nchunk, qb);
    const long long ge = p->g_elems, be = p->big_elems;
    bool ok = dalloc(p, &p->d_tw, (size_t)(4096 + p->maxF / 4096)) && dalloc(p, &p->d_S, n) &&
              dalloc(p, &p->d_perm, n) && dalloc(p, &p->d_payout, k) &&
              dalloc(p, &p->d_blocks, (size_t)qb * g.nb * (B + 1)) && dalloc(p, &p->d_g[0], (size_t)qb * ge) &&
              dalloc(p, &p->d_g[1], (size_t)qb * ge) && dalloc(p, &p->d_partial, (size_t)p->nchunk * n) &&
              dalloc(p, &p->d_acc, n) && dalloc(p, &p->d_equity, n);
    if (ok && be > 0) ok = dalloc(p, &p->d_bufA, (size_t)qb * be) && dalloc(p, &p->d_bufB, (size_t)qb * be);
    p->d_coef.assign(g.L, nullptr);
    for (int l = 1; ok && l <= g.L - 2; ++l) ok = dalloc(p, &p->d_coef[l], (size_t)qb * g.n_real[l] * g.c[l]);
    if (ok && spec) ok = spec_alloc(p);
    if (ok && p->small_fits) ok = dalloc(p, &p->d_small, (size_t)p->small_gmax * n);
    return ok;
}

// Plan creation.  dry: host-only planning against a simulated device with
// dry_free bytes free (no device call is made and no buffer is uploaded).
IcmHipPlan *plan_create(int n, const double *S, int k, const IcmHipOptions *opts, bool sort_players,
                        int max_q = 256, const IcmHipPlanParams *params = nullptr, bool dry = false,
                        size_t dry_free = 0) {
    if ((!S && !dry) || n < 1 || k < 1 || k > n) { set_err("invalid arguments: need 1 <= k <= n"); return nullptr; }
    for (int i = 0; S && i < n; ++i) {
        if (!(S[i] > 0.0) || !std::isfinite(S[i])) { set_err("stacks must be positive and finite"); return nullptr; }
    }
    PlanCfg cfg;
    if (!plan_config(opts, params, cfg)) return nullptr;
    if (!dry && !ensure_device()) return nullptr;
    IcmHipPlan *p = new IcmHipPlan();
    p->verbose = cfg.verbose;
    p->spec_ok = cfg.spec_ok;
    p->frac = cfg.frac;
    p->ckpt_mode = cfg.ckpt;
    p->mem.dry = dry;
    const int B = pick_B(n, k, cfg.B);
    if (!make_geometry(n, k, B, p->geo)) { set_err("invalid geometry"); delete p; return nullptr; }
    const Geometry &g = p->geo;
    for (int l = 1; l < g.L; ++l) {
        if (g.F[l] > icm::kMaxLdsF && (long long)g.n_real[l] + (g.n_real[l] + 1) / 2 > 65535) {
            set_err("level too wide for the row-pass grid");
            delete p;
            return nullptr;
        }
    }

    if (S) {
        // Stack-descending order, index tie-break (src/gpu/gpu_api.cu:492-502).
        p->perm.resize(n);
        for (int i = 0; i < n; ++i) p->perm[i] = i;
        if (sort_players)
            std::sort(p->perm.begin(), p->perm.end(), [S](int a, int b) {
                return S[a] > S[b] || (S[a] == S[b] && a < b);
            });
        p->S_sorted.resize(n);
        for (int i = 0; i < n; ++i) p->S_sorted[i] = S[p->perm[i]];
        p->Smax = *std::max_element(S, S + n);
    }

    long long ge = 0, be = 0;
    const long long pb = plane_bytes(g, &ge, &be);
    p->g_elems = ge;
    p->big_elems = be;
    size_t freeb = dry_free, totb = 0;
    if (!dry && !HIP_OK(hipMemGetInfo(&freeb, &totb))) { plan_free(p); return nullptr; }
    if (cfg.cap) freeb = std::min(freeb, cfg.cap);
    p->mem.cap = dry ? dry_free : cfg.cap;
    // per-plan arrays that do not scale with the q-batch: stacks, order,
    // payout, accumulators, twiddles and at most nchunk_max q-chunk partials
    const int wgb = (g.nb + 256 / B - 1) / (256 / B);
    const long long nchunk_max = std::min(256, std::max(1, 16384 / std::max(1, wgb)));
    const long long fixed = 8LL * n * (4 + nchunk_max) + 8LL * k + 16LL * (4096 + 4096);
    const long long budget = std::max<long long>(0, (long long)(cfg.frac * (double)freeb) - fixed);
    const int qmax = std::max(1, std::min(max_q, 256));
    // launch-shape limit: row passes put (q-plane x array) on grid.y
    auto shape_cap = [&](int q) {
        for (int l = 1; l < g.L; ++l) {
            if (g.F[l] > icm::kMaxLdsF) {
                const long long arrays = (long long)g.n_real[l] + (g.n_real[l] + 1) / 2;
                while (q > 1 && arrays * q > 65535) --q;
            }
        }
        return q;
    };
    auto planes = [&](long long per_plane) {
        const int q = (int)std::max<long long>(1, budget / std::max<long long>(per_plane, 1));
        return shape_cap(std::min(q, qmax));
    };
    int qb = cfg.qb > 0 ? shape_cap(std::min(cfg.qb, qmax)) : planes(pb);
    // The spectrum cache costs HBM per q-plane; take it while at least
    // min(qb, 32) q-planes still fit in one batch.
    p->spec_elems = spec_plane_elems(g);
    bool spec = false;
    if (p->spec_elems > 0 && cfg.spec_ok) {
        const int qs = planes(pb + 16 * p->spec_elems);
        if (cfg.qb > 0) {
            spec = qs >= qb;
        } else if (qs >= std::min(qb, 32)) {
            qb = qs;
            spec = true;
        }
    }
    p->sb_max = tuned_schoolbook_max(n, k);
    p->maxF = 16;
    for (int l = 1; l < g.L; ++l) p->maxF = std::max(p->maxF, g.F[l]);
    p->maxF = std::max(p->maxF, icm::kMaxLdsF);
    p->twShift = ilog2i(p->maxF);
    // small-instance path: one workgroup per quadrature node (two per CU
    // when the tree fits 78 KB of LDS)
    p->small_fits = small_geometry(n, k, p->sj) <= icm::kSmallLdsB;
    if (p->small_fits) {
        int cus = 256;
        if (!dry && hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g_device) != hipSuccess)
            cus = 256;
        p->small_gmax = std::min(512, std::max(1, cus) * (p->sj.lds_doubles <= icm::kSmallLdsA ? 2 : 1));
        p->small_auto = small_work(p->sj) <= kSmallMaxWork;
    }

    if (!dry && !(HIP_OK(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking)) &&
                  HIP_OK(hipEventCreate(&p->ev0)) && HIP_OK(hipEventCreate(&p->ev1)))) {
        plan_free(p);
        return nullptr;
    }
    // Out-of-memory policy of src/gpu/gpu_memory.cu:267-279: on a failed
    // allocation free everything, halve the q-batch and try again, at most 4
    // times.
    for (p->retries = 0;; ++p->retries) {
        if (alloc_buffers(p, qb, spec)) break;
        free_buffers(p);
        if (qb > 1 && p->retries < 4) {
            qb = std::max(1, qb / 2);
            continue;
        }
        const std::string why = g_err;
        char buf[160];
        snprintf(buf, sizeof buf, "plan allocation failed at q_batch=%d after %d halvings: ", qb, p->retries);
        set_err(buf + why);
        plan_free(p);
        return nullptr;
    }
    if (dry) return p;

    // twiddles, rounded once from long double: 4096-entry coarse table
    // exp(-2 pi i m / 4096) followed by the maxF/4096-entry fine table
    // exp(-2 pi i f / maxF) (layout documented in fft_lds.hpp)
    const int nfine = p->maxF / 4096;
    std::vector<double2> tw(4096 + nfine);
    const long double PI = 3.141592653589793238462643383279502884L;
    for (int m = 0; m < 4096; ++m) {
        long double ang = -2.0L * PI * (long double)m / 4096.0L;
        tw[m] = make_double2((double)cosl(ang), (double)sinl(ang));
    }
    for (int f = 0; f < nfine; ++f) {
        long double ang = -2.0L * PI * (long double)f / (long double)p->maxF;
        tw[4096 + f] = make_double2((double)cosl(ang), (double)sinl(ang));
    }
    const bool ok = HIP_OK(hipMemcpy(p->d_tw, tw.data(), sizeof(double2) * tw.size(), hipMemcpyHostToDevice)) &&
                    HIP_OK(hipMemcpy(p->d_S, p->S_sorted.data(), sizeof(double) * n, hipMemcpyHostToDevice)) &&
                    HIP_OK(hipMemcpy(p->d_perm, p->perm.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    if (!ok) { plan_free(p); return nullptr; }

    if (p->verbose) {
        fprintf(stderr, "[icm_hip] n=%d k=%d B=%d nb=%d L=%d qb=%d nchunk=%d spec=%d retries=%d bytes=%.3f GB\n", n,
                k, B, g.nb, g.L, p->qb, p->nchunk, p->d_spec ? 1 : 0, p->retries, p->mem.bytes / 1e9);
        for (int l = 0; l < g.L; ++l)
            fprintf(stderr, "  l=%2d n_real=%8d c=%8d G=%8d F=%8d\n", l, g.n_real[l], g.c[l], g.G[l], g.F[l]);
    }
    return p;
}

bool ensure_nodes(IcmHipPlan *p, int Qa) {
    if (Qa <= p->q_cap) return true;
    if (p->d_nodes) {
        (void)hipFree(p->d_nodes);
        p->d_nodes = nullptr;
        p->mem.bytes -= sizeof(double) * 2 * (size_t)p->q_cap;
    }
    p->q_cap = 0;
    p->nodes_Q = 0;
    if (!dalloc(p, &p->d_nodes, (size_t)2 * Qa)) return false;
    p->q_cap = Qa;
    return true;
}

// Quadrature nodes for Q, compacted to the active ones (w != 0,
// src/cpu/icm.c:2558): logv at d_nodes[0 .. Qa), w at d_nodes[q_cap ..).
// Uploaded once per Q (they depend on Q and max S only); the raw per-node
// entry uses one node with logv = 0, w = 1.  Returns Qa, or -1.
int upload_nodes(IcmHipPlan *p, int Q) {
    const int key = p->raw_node ? -1 : Q;
    if (p->nodes_Q == key && p->d_nodes) return p->nodes_Qa;
    std::vector<double> logv(Q), w(Q);
    make_nodes(Q, p->Smax, logv.data(), w.data());
    if (p->raw_node) { logv.assign(1, 0.0); w.assign(1, 1.0); Q = 1; }
    std::vector<double> lv_a, w_a;
    for (int q = 0; q < Q; ++q)
        if (w[q] != 0.0) { lv_a.push_back(logv[q]); w_a.push_back(w[q]); }
    const int Qa = (int)lv_a.size();
    const int Qs = std::max(Qa, 1);
    if (!ensure_nodes(p, Qs)) return -1;
    std::vector<double> nodes(2 * (size_t)p->q_cap, 0.0);
    std::copy(lv_a.begin(), lv_a.end(), nodes.begin());
    std::copy(w_a.begin(), w_a.end(), nodes.begin() + p->q_cap);
    if (!HIP_OK(hipMemcpyAsync(p->d_nodes, nodes.data(), sizeof(double) * nodes.size(), hipMemcpyHostToDevice,
                               p->stream)) ||
        !HIP_OK(hipStreamSynchronize(p->stream)))
        return -1;
    p->nodes_Q = key;
    p->nodes_Qa = Qa;
    return Qa;
}

bool use_small(const IcmHipPlan *p) {
    if (!p->small_fits || !p->d_small || p->raw_node || p->d_a_in || p->d_inner) return false;
    return p->small_mode == 1 || (p->small_mode < 0 && p->small_auto);
}

// The whole quadrature of a small instance: k_small over the active nodes,
// then the fixed-order sum over workgroups and the scatter to original order.
int run_small(IcmHipPlan *p, int Qa, const double *d_payout, IcmHipRunStats *stats) {
    hipStream_t s = p->stream;
    icm::SmallJob J = p->sj;
    J.S = p->d_S;
    J.logv = p->d_nodes;
    J.w = p->d_nodes + p->q_cap;
    J.payout = d_payout;
    J.partial = p->d_small;
    J.Qa = Qa;
    const int G = std::max(1, std::min(Qa, p->small_gmax));
    if (!HIP_OK(hipEventRecord(p->ev0, s))) return -1;
    if (Qa > 0 && !hip_ok(icm::launch_small(J, G, s, &p->prof, 2.0 * small_work(J) * Qa), "small-instance kernel"))
        return -1;
    if (!hip_ok(icm::launch_small_finalize(p->d_small, Qa > 0 ? G : 0, J.n, p->d_active, p->d_perm, p->d_acc,
                                           p->d_equity, s, &p->prof),
                "small-instance finalize"))
        return -1;
    if (!HIP_OK(hipEventRecord(p->ev1, s))) return -1;
    if (!HIP_OK(hipStreamSynchronize(s))) return -1;
    if (p->prof.enabled) p->prof.collect();
    if (stats) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, p->ev0, p->ev1);
        memset(stats, 0, sizeof(*stats));
        stats->total_ns = (double)ms * 1e6;
        stats->peak_vram_bytes = p->mem.bytes;
        stats->engine = 1;
        stats->B = 1;
        stats->uncached_fused_levels = J.L - 1;
    }
    return 0;
}

// Level l (1 .. L-1) runs as schoolbook sums when its children keep few
// coefficients (c_{l-1} <= sb_max, within the kernels' LDS stage).
bool use_schoolbook(const IcmHipPlan *p, int l) {
    const Geometry &g = p->geo;
    const int cc = g.c[l - 1];
    return cc <= p->sb_max && cc <= icm::kSbMaxC && icm::sb_stride_build(cc) <= icm::kSbLds &&
           icm::sb_stride_prop(cc, g.G[l - 1]) <= icm::kSbLds;
}

// Diagnostic (icm_hip_plan_set_poison): fills every device buffer that a
// q-batch writes before it reads -- block and level products, g ping-pong,
// four-step scratch, the spectrum caches, the per-chunk partials and, for an
// adjoint batch, the tangent and B-adjoint arrays -- with all-ones bytes (a
// NaN in FP64).  A kernel that reads a position the current batch has not
// written first then turns the result into NaN instead of silently reading a
// value left by an earlier batch or run.
bool poison_batch(IcmHipPlan *p, hipStream_t s, bool vjp) {
    const Geometry &g = p->geo;
    auto fill = [&](void *x, size_t bytes) {
        if (!x || bytes == 0) return true;
        p->poisoned_bytes += (long long)bytes;
        return HIP_OK(hipMemsetAsync(x, 0xFF, bytes, s));
    };
    const size_t qb = (size_t)p->qb;
    bool ok = fill(p->d_blocks, 8 * qb * g.nb * (g.B + 1)) && fill(p->d_g[0], 8 * qb * p->g_elems) &&
              fill(p->d_g[1], 8 * qb * p->g_elems) && fill(p->d_partial, 8 * (size_t)p->nchunk * g.n) &&
              fill(p->d_bufA, 16 * qb * p->big_elems) && fill(p->d_bufB, 16 * qb * p->big_elems) &&
              fill(p->d_spec, p->spec_bytes);
    for (int l = 1; ok && l <= g.L - 2; ++l) ok = fill(p->d_coef[l], 8 * qb * g.n_real[l] * g.c[l]);
    if (!ok || !vjp) return ok;
    const size_t qv = (size_t)p->qv;
    ok = fill(p->d_dblocks, 8 * qv * g.nb * (g.B + 1)) && fill(p->d_bv[0], 8 * qv * p->bv_elems) &&
         fill(p->d_bv[1], 8 * qv * p->bv_elems) && fill(p->d_partial_ds, 8 * (size_t)p->nchunk * g.n);
    for (int l = 1; ok && l <= g.L - 2; ++l) ok = fill(p->d_dcoef[l], 8 * qv * g.n_real[l] * g.c[l]);
    for (size_t i = 0; ok && i < p->d_dseg.size(); ++i) ok = fill(p->d_dseg[i], 8 * qv * p->seg_elems);
    if (ok && p->d_dspec) {
        size_t e = 0;
        for (int l = 1; l <= g.L - 2; ++l)
            if (p->dspec_off[l] >= 0) e += qv * g.n_real[l] * g.F[l];
        ok = fill(p->d_dspec, 16 * e);
    }
    return ok;
}

// Runs the whole quadrature for the active nodes with device payout; leaves
// the result in d_acc (sorted order) and d_equity (original order).
int run_plan(IcmHipPlan *p, int Q, const double *d_payout, IcmHipRunStats *stats) {
    const Geometry &g = p->geo;
    if (Q < 1) { set_err("Q must be positive"); return -1; }
    if (!HIP_OK(hipSetDevice(g_device))) return -1;
    const int Qa = upload_nodes(p, Q);
    if (Qa < 0) return -1;
    if (use_small(p)) return run_small(p, Qa, d_payout, stats);
    p->poisoned_bytes = 0;
    hipStream_t s = p->stream;
    const double *d_lv = p->d_nodes;
    const double *d_w = p->d_nodes + p->q_cap;

    if (!HIP_OK(hipEventRecord(p->ev0, s))) return -1;
    if (!HIP_OK(hipMemsetAsync(p->d_acc, 0, sizeof(double) * g.n, s))) return -1;
    const int B = g.B;
    const int L = g.L;
    int n_lds = 0, n_big = 0;
    for (int l = 1; l < L; ++l) (g.F[l] > icm::kMaxLdsF ? n_big : n_lds)++;

    for (int q0 = 0; q0 < Qa; q0 += p->qb) {
        const int nq = std::min(p->qb, Qa - q0);
        if (p->poison && !poison_batch(p, s, false)) return -1;
        icm::LeafJob lj{};
        lj.S = p->d_S;
        lj.logv = d_lv + q0;
        lj.wq = d_w + q0;
        lj.a_in = p->d_a_in ? p->d_a_in + (long long)q0 * g.n : nullptr;
        lj.blocks = p->d_blocks;
        lj.n = g.n; lj.nb = g.nb; lj.B = B; lj.k = g.k; lj.nq = nq;
        lj.G0 = g.G[0];
        lj.blist = nullptr;
        lj.nbl = g.nb;
        p->prof.tag = 0;
        if (!hip_ok(icm::launch_leaf_build(lj, s, &p->prof), "leaf_build")) return -1;

        // views of level l's coefficients
        auto coef_view = [&](int l) {
            icm::CoefView v{};
            if (l == 0) {
                v.base = p->d_blocks; v.plane = (long long)g.nb * (B + 1); v.stride = B + 1; v.count = g.c[0];
            } else {
                v.base = p->d_coef[l]; v.plane = (long long)g.n_real[l] * g.c[l]; v.stride = g.c[l];
                v.count = g.c[l];
            }
            return v;
        };
        // build; a four-step level whose level above keeps its child-pair
        // spectra in the cache hands them over directly (column fusion: its
        // parent coefficients are then never read; pruned propagates read
        // the cache through their parent lists too)
        std::vector<icm::LevelJob> BJ(L);
        for (int l = 1; l <= L - 2; ++l) {
            icm::LevelJob &J = BJ[l];
            J = icm::LevelJob{};
            J.child = coef_view(l - 1);
            J.parent = icm::CoefOut{p->d_coef[l], (long long)g.n_real[l] * g.c[l], g.c[l], g.c[l]};
            J.nr_child = g.n_real[l - 1];
            J.nr_parent = g.n_real[l];
            J.plist = nullptr;
            J.np = g.n_real[l];
            J.F = g.F[l];
            J.logF = ilog2i(g.F[l]);
            J.nq = nq;
        }
        auto spec_at = [&](int l) -> double2 * {
            return (p->d_spec && l >= 1 && l <= L - 2 && p->spec_off[l] >= 0) ? p->d_spec + p->spec_off[l] : nullptr;
        };
        // the root is not built; its propagate transforms its two children:
        // the top build level may hand them over in bufA instead (the root's
        // child-pair region), which nothing else writes before the root runs
        icm::LevelJob root{};
        root.child = coef_view(L - 2);
        root.nr_child = g.n_real[L - 2];
        root.nr_parent = 1;
        root.np = 1;
        root.F = g.F[L - 1];
        root.logF = ilog2i(g.F[L - 1]);
        root.nq = nq;
        bool child_ready = false, root_ready = false;
        for (int l = 1; l <= L - 2; ++l) {
            const icm::LevelJob &J = BJ[l];
            p->prof.tag = l;
            const bool large = !use_schoolbook(p, l) && J.F > icm::kMaxLdsF;
            const bool top = (l == L - 2);
            const icm::LevelJob &N = top ? root : BJ[l + 1];
            const int bit = top ? 4 : (icm::col_fused(J.F) ? 8 : 2);
            const bool fuse = large && (p->fuse_columns & bit) &&
                              (top ? (L >= 3 && g.F[L - 1] > icm::kMaxLdsF && !use_schoolbook(p, L - 1))
                                   : (spec_at(l + 1) && !use_schoolbook(p, l + 1) && N.F > icm::kMaxLdsF)) &&
                              icm::build_fusable(J, N);
            hipError_t e = use_schoolbook(p, l)
                               ? icm::launch_build_sb(J, s, &p->prof)
                           : (J.F <= icm::kMaxLdsF)
                               ? icm::launch_build_lds(J, p->d_tw, p->twShift, s, &p->prof)
                               : icm::launch_build_large(J, p->d_tw, p->twShift, p->d_bufA, p->d_bufB, s, &p->prof,
                                                         spec_at(l), child_ready, fuse ? &N : nullptr,
                                                         fuse ? (top ? p->d_bufA : spec_at(l + 1)) : nullptr);
            if (!hip_ok(e, "build level")) return -1;
            child_ready = fuse;
            root_ready = top && fuse;
        }
        // propagate: root g = payout (same for every q-plane); jobs of every
        // level first, so that a four-step level can hand the level below its
        // g-pair column transforms directly (column fusion)
        std::vector<icm::LevelJob> PJ(L);
        int cur = 0;
        for (int l = L - 1; l >= 1; --l) {
            icm::LevelJob &J = PJ[l];
            J = icm::LevelJob{};
            J.child = coef_view(l - 1);
            if (l == L - 1) {
                J.gpar = icm::CoefView{d_payout, 0, g.k, g.k};
            } else {
                J.gpar = icm::CoefView{p->d_g[cur], (long long)g.n_real[l] * g.G[l], g.G[l], g.G[l]};
                cur ^= 1;
            }
            J.gchild = icm::CoefOut{p->d_g[cur], (long long)g.n_real[l - 1] * g.G[l - 1], g.G[l - 1], g.G[l - 1]};
            J.nr_child = g.n_real[l - 1];
            J.nr_parent = g.n_real[l];
            if (p->d_lists) {
                J.plist = p->d_lists + p->list_off[l];
                J.np = p->list_cnt[l];
            } else {
                J.plist = nullptr;
                J.np = g.n_real[l];
            }
            J.F = g.F[l];
            J.logF = ilog2i(g.F[l]);
            J.nq = nq;
        }
        bool g_ready = false;
        for (int l = L - 1; l >= 1; --l) {
            const icm::LevelJob &J = PJ[l];
            if (J.np == 0) { g_ready = false; continue; }
            p->prof.tag = -l;
            const bool large = !use_schoolbook(p, l) && J.F > icm::kMaxLdsF;
            const bool fuse = large && (p->fuse_columns & 1) && l >= 2 && !use_schoolbook(p, l - 1) &&
                              PJ[l - 1].F > icm::kMaxLdsF && icm::prop_fusable(J, PJ[l - 1]);
            hipError_t e = use_schoolbook(p, l)
                               ? icm::launch_prop_sb(J, s, &p->prof)
                           : (J.F <= icm::kMaxLdsF)
                               ? icm::launch_prop_lds(J, p->d_tw, p->twShift, s, &p->prof)
                               : icm::launch_prop_large(J, p->d_tw, p->twShift, p->d_bufA, p->d_bufB, s, &p->prof,
                                                        (l == L - 1 && root_ready)
                                                            ? p->d_bufA
                                                        : (p->d_spec && l <= L - 2 && p->spec_off[l] >= 0)
                                                            ? p->d_spec + p->spec_off[l]
                                                            : nullptr,
                                                        g_ready, fuse ? &PJ[l - 1] : nullptr);
            if (!hip_ok(e, "propagate level")) return -1;
            g_ready = fuse;
        }
        if (L == 1) {
            lj.g = d_payout; lj.g_plane = 0; lj.g_stride = g.k;
        } else {
            lj.g = p->d_g[cur]; lj.g_plane = (long long)g.n_real[0] * g.G[0]; lj.g_stride = g.G[0];
        }
        const int nchunk = std::min(p->nchunk, nq);
        lj.q_per_chunk = (nq + nchunk - 1) / nchunk;
        const int nch = (nq + lj.q_per_chunk - 1) / lj.q_per_chunk;
        lj.partial = p->d_partial;
        lj.inner_out = p->d_inner ? p->d_inner + (long long)q0 * g.n : nullptr;
        lj.active = p->d_active;
        p->prof.tag = 0;
        if (p->d_lists) {
            lj.blist = p->d_lists + p->list_off[L];
            lj.nbl = p->list_cnt[L];
        }
        if (lj.nbl > 0 && !hip_ok(icm::launch_leaf_extract(lj, nch, s, &p->prof), "leaf_extract")) return -1;
        if (!hip_ok(icm::launch_finalize(p->d_partial, nch, g.n, p->d_active, p->d_acc, 1, s, &p->prof),
                    "finalize"))
            return -1;
    }
    if (!hip_ok(icm::launch_scatter(p->d_acc, g.n, p->d_perm, p->d_equity, s, &p->prof), "scatter")) return -1;
    if (!HIP_OK(hipEventRecord(p->ev1, s))) return -1;
    if (!HIP_OK(hipStreamSynchronize(s))) return -1;
    if (p->prof.enabled) p->prof.collect();
    if (stats) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, p->ev0, p->ev1);
        memset(stats, 0, sizeof(*stats));
        stats->total_ns = (double)ms * 1e6;
        stats->peak_vram_bytes = p->mem.bytes;
        stats->engine = 1;
        stats->B = B;
        stats->uncached_fused_levels = n_lds;
        stats->uncached_cufft_levels = n_big;
    }
    return 0;
}

// Frees the adjoint buffers (a failed or partial allocation leaves none).
void free_vjp(IcmHipPlan *p, size_t bytes_before) {
    auto f = [](auto *&x) { if (x) (void)hipFree((void *)x); x = nullptr; };
    f(p->d_dblocks); for (double *&c : p->d_dcoef) f(c);
    for (double *&c : p->d_dseg) f(c);
    f(p->d_bv[0]); f(p->d_bv[1]); f(p->d_dspec); f(p->d_v); f(p->d_partial_ds); f(p->d_acc_ds); f(p->d_ds);
    p->qv = 0;
    p->ckpt = 0;
    p->mem.bytes = bytes_before;
}

// Tangent level l's array under checkpoint interval m: its own for every
// level when m < 2 and for the checkpoint levels l = m, 2m, ...; otherwise
// slot (l mod m) - 1, which holds one segment's levels at a time.
static bool ckpt_keeps(int m, int l) { return m < 2 || l % m == 0; }
static double *dcoef_of(IcmHipPlan *p, int l) {
    return ckpt_keeps(p->ckpt, l) ? p->d_dcoef[l] : p->d_dseg[l % p->ckpt - 1];
}

// VJP buffers: tangent block/level products, B adjoint ping-pong, dS arrays.
// Sized from the plan's memory share of what is free now.  Checkpointing
// (interval m) keeps the tangent products of levels 0, m, 2m, ... plus m - 1
// slots for one segment, recomputed from its checkpoint in the down-sweep;
// auto mode (ckpt_mode -1) takes the interval that leaves the most q-planes
// when keeping every level leaves fewer than min(q-batch, 32).  On a failed
// allocation the q-planes are halved, at most 4 times, as at plan creation.
bool ensure_vjp(IcmHipPlan *p) {
    if (p->qv > 0) return true;
    const Geometry &g = p->geo;
    long long be = 0;
    for (int l = 0; l < g.L; ++l) be = std::max(be, (long long)g.n_real[l] * g.c[l]);
    auto lev = [&](int l) { return (long long)g.n_real[l] * g.c[l]; };
    auto seg_of = [&](int m) {       // elements per slot
        long long e = 0;
        for (int l = 1; l <= g.L - 2; ++l)
            if (!ckpt_keeps(m, l)) e = std::max(e, lev(l));
        return e;
    };
    auto per_of = [&](int m) {       // bytes per q-plane
        long long e = (long long)g.nb * (g.B + 1) + 2 * be;
        for (int l = 1; l <= g.L - 2; ++l)
            if (ckpt_keeps(m, l)) e += lev(l);
        if (m >= 2) e += (long long)(m - 1) * seg_of(m);
        return 8 * e;
    };
    size_t freeb = 0, totb = 0;
    if (!HIP_OK(hipMemGetInfo(&freeb, &totb))) return false;
    if (p->mem.cap) freeb = std::min(freeb, p->mem.cap - std::min(p->mem.cap, p->mem.bytes));
    const long long budget = (long long)(p->frac * (double)freeb);
    auto planes = [&](long long per) {
        return (int)std::min<long long>(p->qb, std::max<long long>(1, budget / std::max<long long>(per, 1)));
    };
    int m = p->ckpt_mode >= 2 ? p->ckpt_mode : 0;
    int qv = planes(per_of(m));
    if (p->ckpt_mode == -1 && qv < std::min(p->qb, 32)) {
        int best = 0;
        long long bper = per_of(0);
        for (int mm = 2; mm <= std::max(2, g.L - 2); ++mm)
            if (per_of(mm) < bper) { bper = per_of(mm); best = mm; }
        if (best && planes(bper) > qv) { m = best; qv = planes(bper); }
    }
    // the tangent spectrum cache, beside the forward's, while the q-batch
    // keeps at least min(qv, 32) planes; also with checkpointing: a tangent
    // build writes its level's entry whenever it runs (recomputed segments
    // rewrite the same values), and with it the four-step levels' down-sweep
    // reads only spectra, so their tangent products need no recompute
    bool dspec = false;
    if (p->d_spec && p->spec_elems > 0) {
        const int qd = planes(per_of(m) + 16 * p->spec_elems);
        if (qd >= std::min(qv, 32)) {
            qv = qd;
            dspec = true;
        }
    }
    p->bv_elems = be;
    p->seg_elems = m >= 2 ? seg_of(m) : 0;
    const size_t before = p->mem.bytes;
    for (int attempt = 0;; ++attempt) {
        p->ckpt = m;
        bool ok = dalloc(p, &p->d_dblocks, (size_t)qv * g.nb * (g.B + 1)) &&
                  dalloc(p, &p->d_bv[0], (size_t)qv * be) && dalloc(p, &p->d_bv[1], (size_t)qv * be) &&
                  dalloc(p, &p->d_v, g.n) && dalloc(p, &p->d_partial_ds, (size_t)p->nchunk * g.n) &&
                  dalloc(p, &p->d_acc_ds, g.n) && dalloc(p, &p->d_ds, g.n);
        p->d_dcoef.assign(g.L, nullptr);
        for (int l = 1; ok && l <= g.L - 2; ++l)
            if (ckpt_keeps(m, l)) ok = dalloc(p, &p->d_dcoef[l], (size_t)qv * lev(l));
        p->d_dseg.assign(m >= 2 ? m - 1 : 0, nullptr);
        for (size_t i = 0; ok && i < p->d_dseg.size(); ++i)
            ok = p->seg_elems == 0 || dalloc(p, &p->d_dseg[i], (size_t)qv * p->seg_elems);
        if (ok && dspec) {
            p->dspec_off.assign(g.L, -1);
            long long off = 0;
            for (int l = 1; l <= g.L - 2; ++l) {
                if (spec_level(g, l)) {
                    p->dspec_off[l] = off;
                    off += (long long)qv * g.n_real[l] * g.F[l];
                }
            }
            // without the tangent cache the tangent builds take three passes
            if (off > 0 && !dalloc(p, &p->d_dspec, (size_t)off)) p->d_dspec = nullptr;
        }
        if (ok) break;
        free_vjp(p, before);
        if (qv > 1 && attempt < 4) {
            qv = std::max(1, qv / 2);
            continue;
        }
        const std::string why = g_err;
        set_err("adjoint buffers: " + why);
        return false;
    }
    p->qv = qv;
    p->vjp_base = before;
    return true;
}

// Adjoint pass: equity and dS = d(v.E)/dS (quadrature nodes held fixed).
// d_v: cotangent in original order (device).  Leaves d_equity and d_ds.
int run_vjp(IcmHipPlan *p, int Q, const double *d_payout, const double *d_v_orig, IcmHipRunStats *stats) {
    const Geometry &g = p->geo;
    if (Q < 1) { set_err("Q must be positive"); return -1; }
    if (!HIP_OK(hipSetDevice(g_device))) return -1;
    if (!ensure_vjp(p)) return -1;
    const int Qa = upload_nodes(p, Q);
    if (Qa < 0) return -1;
    hipStream_t s = p->stream;
    p->poisoned_bytes = 0;
    const double *d_lv = p->d_nodes, *d_w = p->d_nodes + p->q_cap;
    if (!HIP_OK(hipEventRecord(p->ev0, s))) return -1;
    if (d_v_orig) {
        if (!hip_ok(icm::launch_gather(d_v_orig, p->d_perm, g.n, p->d_v, s), "gather")) return -1;
    } else if (!HIP_OK(hipMemsetAsync(p->d_v, 0, sizeof(double) * g.n, s))) {
        return -1;
    }
    if (!HIP_OK(hipMemsetAsync(p->d_acc, 0, sizeof(double) * g.n, s))) return -1;
    if (!HIP_OK(hipMemsetAsync(p->d_acc_ds, 0, sizeof(double) * g.n, s))) return -1;
    const int B = g.B, L = g.L;
    auto view = [&](int l, bool tangent) {
        icm::CoefView v{};
        if (l == 0) {
            v.base = tangent ? p->d_dblocks : p->d_blocks;
            v.plane = (long long)g.nb * (B + 1); v.stride = B + 1; v.count = g.c[0];
        } else {
            v.base = tangent ? dcoef_of(p, l) : p->d_coef[l];
            v.plane = (long long)g.n_real[l] * g.c[l]; v.stride = g.c[l]; v.count = g.c[l];
        }
        return v;
    };
    // spec: the level's spectrum cache, for the passes whose child pairs are
    // the forward products P (built once, read by the A and B propagates)
    auto level = [&](icm::LevelJob &J, bool build, double2 *spec = nullptr, bool g_ready = false) -> bool {
        hipError_t e;
        if (build)
            e = (J.F <= icm::kMaxLdsF) ? icm::launch_build_lds(J, p->d_tw, p->twShift, s, &p->prof)
                                       : icm::launch_build_large(J, p->d_tw, p->twShift, p->d_bufA, p->d_bufB, s,
                                                                 &p->prof, spec);
        else
            e = (J.F <= icm::kMaxLdsF) ? icm::launch_prop_lds(J, p->d_tw, p->twShift, s, &p->prof)
                                       : icm::launch_prop_large(J, p->d_tw, p->twShift, p->d_bufA, p->d_bufB, s,
                                                                &p->prof, spec, g_ready);
        return hip_ok(e, build ? "vjp build level" : "vjp propagate level");
    };
    auto spec_of = [&](int l) -> double2 * {
        return (p->d_spec && spec_level(g, l)) ? p->d_spec + p->spec_off[l] : nullptr;
    };
    auto dspec_of = [&](int l) -> double2 * {
        return (p->d_dspec && spec_of(l)) ? p->d_dspec + p->dspec_off[l] : nullptr;
    };
    const int qstep = std::min(p->qb, p->qv);
    for (int q0 = 0; q0 < Qa; q0 += qstep) {
        const int nq = std::min(qstep, Qa - q0);
        if (p->poison && !poison_batch(p, s, true)) return -1;
        icm::VjpLeafJob V{};
        icm::LeafJob &lj = V.L;
        lj.S = p->d_S; lj.logv = d_lv + q0; lj.wq = d_w + q0;
        lj.a_in = p->d_a_in ? p->d_a_in + (long long)q0 * g.n : nullptr;
        lj.blocks = p->d_blocks;
        lj.n = g.n; lj.nb = g.nb; lj.B = B; lj.k = g.k; lj.nq = nq; lj.G0 = g.G[0];
        lj.blist = nullptr; lj.nbl = g.nb;
        V.v = p->d_v;
        V.c_in = p->d_c_in ? p->d_c_in + (long long)q0 * g.n : nullptr;
        V.dblocks = p->d_dblocks;
        V.c0 = g.c[0];
        if (!hip_ok(icm::launch_leaf_build_vjp(V, s, &p->prof), "leaf_build_vjp")) return -1;
        // P_X and dP_X of level l (recompute: dP_X only; P is kept for every
        // level, and a fused LDS level rewrites it with the same values)
        auto pjob = [&](int l) {   // P_X of level l from its children's P
            icm::LevelJob J{};
            J.nr_child = g.n_real[l - 1]; J.nr_parent = g.n_real[l]; J.np = g.n_real[l];
            J.F = g.F[l]; J.logF = ilog2i(g.F[l]); J.nq = nq;
            J.child = view(l - 1, false);
            J.parent = icm::CoefOut{p->d_coef[l], (long long)g.n_real[l] * g.c[l], g.c[l], g.c[l]};
            return J;
        };
        // Up-sweep column fusion: a four-step P build hands the level above
        // its child-pair column transforms (the spectrum cache entry) when
        // every later reader of this level's P takes those spectra, i.e. the
        // level above keeps both spectrum caches (its dP build and its A and
        // B propagates then read spectra, and the exact terms read the end
        // coefficients, which the fused kernels store).  p_ready: the level
        // below handed this level's entry over.
        // d_ready: likewise for the tangent builds (the level below's dP
        // column inverse wrote this level's dP-pair transforms into its cache
        // entry).  partial[l]: level l's P or dP coefficients were handed over
        // and only their end coefficients stored; nothing may read them whole
        // (a rebuild from them is refused -- the fusions are chosen so that
        // none is needed: a level with both caches is never rebuilt).
        bool d_ready = false;
        std::vector<char> partial(L + 1, 0);
        bool p_ready = false;
        auto build_level = [&](int l, bool recompute) -> bool {
            icm::LevelJob J = pjob(l);
            if (J.F <= icm::kMaxLdsF) {                           // P_X and dP_X in one pass
                J.tchild = view(l - 1, true);
                J.tparent = icm::CoefOut{dcoef_of(p, l), J.parent.plane, J.parent.stride, J.parent.count};
                p_ready = false;
                d_ready = false;
                return hip_ok(icm::launch_build_lds_t(J, p->d_tw, p->twShift, s, &p->prof), "vjp build level");
            }
            if (!recompute) {                                     // P_X = P_L P_R
                icm::LevelJob N{};
                bool fuse = false;
                if (l + 1 <= L - 2 && spec_of(l + 1) && dspec_of(l + 1) &&
                    (p->fuse_columns & (icm::col_fused(J.F) ? 8 : 2))) {
                    N = pjob(l + 1);
                    fuse = icm::build_fusable(J, N);
                }
                if (!hip_ok(icm::launch_build_large(J, p->d_tw, p->twShift, p->d_bufA, p->d_bufB, s, &p->prof,
                                                    spec_of(l), p_ready && spec_of(l), fuse ? &N : nullptr,
                                                    fuse ? spec_of(l + 1) : nullptr),
                            "vjp build level"))
                    return false;
                p_ready = fuse;
                if (fuse) partial[l] = 1;
            }
            J.parent.base = dcoef_of(p, l);
            if (recompute && partial[l - 1]) {
                set_err("adjoint checkpoint: a rebuilt level would read coefficients handed over as spectra");
                return false;
            }
            if (dspec_of(l)) {                                    // dP_X from both pairs' spectra
                J.tchild = view(l - 1, true);
                icm::LevelJob N{};
                bool fuse = false;
                if (!recompute && l + 1 <= L - 2 && spec_of(l + 1) && dspec_of(l + 1) && (p->fuse_columns & 2)) {
                    N = pjob(l + 1);
                    N.tchild = view(l, true);
                    fuse = icm::build_fusable_t(J, N);
                }
                if (!hip_ok(icm::launch_build_large_t(J, p->d_tw, p->twShift, spec_of(l), dspec_of(l), p->d_bufB, s,
                                                      &p->prof, d_ready && !recompute, fuse ? &N : nullptr,
                                                      fuse ? dspec_of(l + 1) : nullptr),
                            "vjp tangent build level"))
                    return false;
                if (!recompute) {
                    d_ready = fuse;
                    if (fuse) partial[l] = 1;
                }
                return true;
            }
            if (!recompute) d_ready = false;
            J.child = view(l - 1, true);
            J.child_r = view(l - 1, false);
            if (!level(J, true)) return false;                    // dP_X = dP_L P_R
            J.child = view(l - 1, false);
            J.child_r = view(l - 1, true);
            J.r_missing_zero = 1;
            J.acc_out = 1;
            return level(J, true);                                //      + P_L dP_R
        };
        for (int l = 1; l <= L - 2; ++l)
            if (!build_level(l, false)) return -1;
        // checkpointing: the slots now hold the segment of the highest
        // non-checkpoint level; a lower segment is rebuilt from its
        // checkpoint when the down-sweep first needs one of its levels
        const int m = p->ckpt;
        int resident = -1;
        for (int l = L - 2; m >= 2 && l >= 1; --l)
            if (!ckpt_keeps(m, l)) { resident = l / m; break; }
        int ca = 0, cb = 0;  // current A / B buffers (A in d_g, B in d_bv)
        // Down-sweep column fusion: a four-step level whose B propagate takes
        // the one-pass form (both spectrum caches) runs it before its A
        // propagate, whose column inverse may then write the level below's A
        // g-pair column transforms straight into bufA (where this level's
        // were) when the level below takes the same form; that level's A
        // coefficients are then never stored except coefficient 0, which its
        // B propagate's wrapped term reads.  a_ready: bufA holds this level's.
        // b_ready: likewise for B (the level above's B column inverse wrote
        // this level's B g-pair transforms; B's coefficients 0 and F / 2 kept).
        bool a_ready = false, b_ready = false;
        auto one_pass_b = [&](int l) {
            return l < L - 1 && g.F[l] > icm::kMaxLdsF && !icm::col_fused(g.F[l]) && dspec_of(l) != nullptr;
        };
        for (int l = L - 1; l >= 1; --l) {
            const bool root = (l == L - 1);
            // this level reads the children's tangent products as coefficients
            // (LDS levels, or four-step levels without the tangent spectra)
            const bool needs_dp = g.F[l] <= icm::kMaxLdsF || !dspec_of(l) || root;
            if (m >= 2 && needs_dp && l - 1 >= 1 && !ckpt_keeps(m, l - 1) && (l - 1) / m != resident) {
                for (int j = (l - 1) / m * m + 1; j <= l - 1; ++j)
                    if (!build_level(j, true)) return -1;
                resident = (l - 1) / m;
            }
            icm::CoefView Apar = root ? icm::CoefView{d_payout, 0, g.k, g.k}
                                      : icm::CoefView{p->d_g[ca], (long long)g.n_real[l] * g.G[l], g.G[l], g.G[l]};
            icm::CoefView Bpar{p->d_bv[cb], (long long)g.n_real[l] * g.c[l], g.c[l], g.c[l]};
            const int na = root ? 0 : ca ^ 1, nbv = root ? 0 : cb ^ 1;
            icm::LevelJob J{};
            J.nr_child = g.n_real[l - 1]; J.nr_parent = g.n_real[l]; J.np = g.n_real[l];
            J.F = g.F[l]; J.logF = ilog2i(g.F[l]); J.nq = nq;
            J.child = view(l - 1, false);
            J.gpar = Apar;
            J.gchild = icm::CoefOut{p->d_g[na], (long long)g.n_real[l - 1] * g.G[l - 1], g.G[l - 1], g.G[l - 1]};
            const int cc = g.c[l - 1];
            icm::CoefOut Bout{p->d_bv[nbv], (long long)g.n_real[l - 1] * cc, cc, cc};
            const int gfix = (g.F[l] < 2 * cc - 1) ? 1 : 0;
            if (J.F <= icm::kMaxLdsF) {                           // A and B adjoints in one pass
                J.tchild = view(l - 1, true);
                J.bpar = root ? icm::CoefView{} : Bpar;
                J.bchild = Bout;
                J.gfix = gfix;
                if (!hip_ok(icm::launch_prop_lds_t(J, p->d_tw, p->twShift, s, &p->prof), "vjp propagate level"))
                    return -1;
                ca = na;
                cb = nbv;
                continue;
            }
            icm::LevelJob K = J;                                  // B_L = A_X * dR ...
            K.child = view(l - 1, true);
            K.r_missing_zero = 1;
            K.gchild = Bout;
            K.gfix = gfix;
            if (one_pass_b(l)) {                                  // ... + B_X * R in one pass, then A
                icm::LevelJob KB = J;
                KB.gpar = Bpar;
                KB.gchild = Bout;
                KB.gfix = gfix;
                if (!a_ready && !hip_ok(icm::launch_g_pairs(J, p->d_tw, p->twShift, p->d_bufA, s, &p->prof),
                                        "vjp propagate level"))
                    return -1;
                icm::LevelJob NB{};                               // the level below's B propagate (B_X * R)
                bool fuse_b = false;
                if (l - 1 >= 1 && one_pass_b(l - 1) && (p->fuse_columns & 1)) {
                    NB.nr_child = g.n_real[l - 2]; NB.nr_parent = g.n_real[l - 1]; NB.np = g.n_real[l - 1];
                    NB.F = g.F[l - 1]; NB.logF = ilog2i(g.F[l - 1]); NB.nq = nq;
                    NB.child = view(l - 2, false);
                    NB.gpar = icm::CoefView{p->d_bv[nbv], Bout.plane, Bout.stride, Bout.count};
                    fuse_b = icm::prop_b_fusable(K, NB);
                }
                if (!hip_ok(icm::launch_prop_large_b(K, KB, p->d_tw, p->twShift, p->d_bufA, p->d_bufB, spec_of(l),
                                                     dspec_of(l), s, &p->prof, b_ready, fuse_b ? &NB : nullptr),
                            "vjp propagate level"))
                    return -1;
                b_ready = fuse_b;
                icm::LevelJob N{};                                // the level below's A propagate
                bool fuse = false;
                if (l - 1 >= 1 && one_pass_b(l - 1) && (p->fuse_columns & 1)) {
                    N.nr_child = g.n_real[l - 2]; N.nr_parent = g.n_real[l - 1]; N.np = g.n_real[l - 1];
                    N.F = g.F[l - 1]; N.logF = ilog2i(g.F[l - 1]); N.nq = nq;
                    N.child = view(l - 2, false);
                    N.gpar = icm::CoefView{p->d_g[na], (long long)g.n_real[l - 1] * g.G[l - 1], g.G[l - 1], g.G[l - 1]};
                    N.gchild = icm::CoefOut{p->d_g[na ^ 1], (long long)g.n_real[l - 2] * g.G[l - 2], g.G[l - 2],
                                            g.G[l - 2]};
                    // the wrapped term reads coefficients 0 and F of the level below's A; only 0 is kept
                    fuse = icm::prop_fusable(J, N) && N.gpar.count <= N.F;
                }
                if (!hip_ok(icm::launch_prop_large(J, p->d_tw, p->twShift, p->d_bufA, p->d_bufB, s, &p->prof,
                                                   spec_of(l), true, fuse ? &N : nullptr, fuse),
                            "vjp propagate level"))   // A_L = A_X * R
                    return -1;
                a_ready = fuse;
                ca = na;
                cb = nbv;
                continue;
            }
            a_ready = b_ready = false;
            if (!level(J, false, spec_of(l))) return -1;         // A_L = A_X * R
            if (icm::col_fused(J.F) && !root) {                  // F = 8192: B in one pass from the coefficients
                icm::LevelJob KB = J;
                KB.gpar = Bpar;
                KB.gchild = Bout;
                KB.gfix = gfix;
                if (!hip_ok(icm::launch_prop_large_b(K, KB, p->d_tw, p->twShift, p->d_bufA, p->d_bufB, nullptr,
                                                     nullptr, s, &p->prof),
                            "vjp propagate level"))
                    return -1;
                ca = na;
                cb = nbv;
                continue;
            }
            // same g pair as the pass above: its column transforms are still in bufA
            if (!level(K, false, dspec_of(l), true)) return -1;
            if (!root) {
                icm::LevelJob K = J;
                K.gpar = Bpar;
                K.gchild = Bout;
                K.gfix = gfix;
                K.acc_out = 1;
                if (!level(K, false, spec_of(l))) return -1;     //     + B_X * R
            }
            ca = na;
            cb = nbv;
        }
        if (L == 1) {
            lj.g = d_payout; lj.g_plane = 0; lj.g_stride = g.k;
            V.bvec = nullptr;
        } else {
            lj.g = p->d_g[ca]; lj.g_plane = (long long)g.n_real[0] * g.G[0]; lj.g_stride = g.G[0];
            V.bvec = p->d_bv[cb]; V.b_plane = (long long)g.n_real[0] * g.c[0]; V.b_stride = g.c[0];
        }
        const int nchunk = std::min(p->nchunk, nq);
        lj.q_per_chunk = (nq + nchunk - 1) / nchunk;
        const int nch = (nq + lj.q_per_chunk - 1) / lj.q_per_chunk;
        lj.partial = p->d_partial;
        lj.inner_out = p->d_inner ? p->d_inner + (long long)q0 * g.n : nullptr;
        V.partial_ds = p->d_partial_ds;
        V.h_out = p->d_h ? p->d_h + (long long)q0 * g.n : nullptr;
        if (!hip_ok(icm::launch_leaf_vjp(V, nch, s, &p->prof), "leaf_vjp")) return -1;
        if (!hip_ok(icm::launch_finalize(p->d_partial, nch, g.n, nullptr, p->d_acc, 1, s, &p->prof), "finalize") ||
            !hip_ok(icm::launch_finalize(p->d_partial_ds, nch, g.n, nullptr, p->d_acc_ds, 1, s, &p->prof),
                    "finalize"))
            return -1;
    }
    if (!hip_ok(icm::launch_scatter(p->d_acc, g.n, p->d_perm, p->d_equity, s, &p->prof), "scatter") ||
        !hip_ok(icm::launch_scatter(p->d_acc_ds, g.n, p->d_perm, p->d_ds, s, &p->prof), "scatter"))
        return -1;
    if (!HIP_OK(hipEventRecord(p->ev1, s))) return -1;
    if (!HIP_OK(hipStreamSynchronize(s))) return -1;
    if (p->prof.enabled) p->prof.collect();
    if (stats) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, p->ev0, p->ev1);
        memset(stats, 0, sizeof(*stats));
        stats->total_ns = (double)ms * 1e6;
        stats->peak_vram_bytes = p->mem.bytes;
        stats->engine = 1;
        stats->B = B;
    }
    return 0;
}

int equity_host(IcmHipPlan *p, int Q, const double *payout, double *equity, IcmHipRunStats *stats) {
    if (!p || !payout || !equity) { set_err("null argument"); return -1; }
    if (!HIP_OK(hipSetDevice(g_device))) return -1;
    if (!HIP_OK(hipMemcpyAsync(p->d_payout, payout, sizeof(double) * p->geo.k, hipMemcpyHostToDevice,
                               p->stream)))
        return -1;
    if (run_plan(p, Q, p->d_payout, stats) != 0) return -1;
    if (!HIP_OK(hipMemcpy(equity, p->d_equity, sizeof(double) * p->geo.n, hipMemcpyDeviceToHost))) return -1;
    return 0;
}

// Subset: active mask (sorted order) plus the pruning lists of
// src/cpu/icm.c:2430-2473 restated for this geometry: block b is hot if it
// holds a target; a node is hot if either child is; a propagate level l runs
// only over parents with a hot child, and the leaf extract only over hot
// blocks.  The build is unchanged (every product feeds its parent).
int set_targets(IcmHipPlan *p, const int *targets, int n_targets) {
    const Geometry &g = p->geo;
    const int n = g.n;
    std::vector<int> inv(n);
    for (int i = 0; i < n; ++i) inv[p->perm[i]] = i;
    std::vector<unsigned char> act(n, 0);
    for (int t = 0; t < n_targets; ++t)
        if (targets[t] >= 0 && targets[t] < n) act[inv[targets[t]]] = 1;  // out-of-range ignored (gpu_api.cu:636-641)
    std::vector<std::vector<unsigned char>> hot(g.L);
    hot[0].assign(g.nb, 0);
    for (int i = 0; i < n; ++i)
        if (act[i]) hot[0][i / g.B] = 1;
    for (int l = 1; l < g.L; ++l) {
        hot[l].assign(g.n_real[l], 0);
        for (int j = 0; j < g.n_real[l]; ++j) {
            const bool r = 2 * j + 1 < g.n_real[l - 1] && hot[l - 1][2 * j + 1];
            hot[l][j] = (unsigned char)(hot[l - 1][2 * j] || r);
        }
    }
    std::vector<int> lists;
    p->list_off.assign(g.L + 1, 0);
    p->list_cnt.assign(g.L + 1, 0);
    for (int l = 1; l < g.L; ++l) {
        p->list_off[l] = (int)lists.size();
        for (int j = 0; j < g.n_real[l]; ++j)
            if (hot[l][j]) lists.push_back(j);
        p->list_cnt[l] = (int)lists.size() - p->list_off[l];
    }
    p->list_off[g.L] = (int)lists.size();
    for (int b = 0; b < g.nb; ++b)
        if (hot[0][b]) lists.push_back(b);
    p->list_cnt[g.L] = (int)lists.size() - p->list_off[g.L];
    if (!p->d_active && !dalloc(p, &p->d_active, n)) return -1;
    if (p->d_lists) { (void)hipFree(p->d_lists); p->d_lists = nullptr; }
    if (!dalloc(p, &p->d_lists, std::max<size_t>(1, lists.size()))) return -1;
    return (HIP_OK(hipMemcpy(p->d_active, act.data(), n, hipMemcpyHostToDevice)) &&
            (lists.empty() ||
             HIP_OK(hipMemcpy(p->d_lists, lists.data(), sizeof(int) * lists.size(), hipMemcpyHostToDevice))))
               ? 0
               : -1;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int icm_hip_init(int device_id) {
    std::lock_guard<std::mutex> lk(g_mu);
    int cnt = 0;
    if (!HIP_OK(hipGetDeviceCount(&cnt))) return 0;
    if (device_id < 0 || device_id >= cnt) { set_err("device id out of range"); return 0; }
    if (!HIP_OK(hipSetDevice(device_id))) return 0;
    g_device = device_id;
    return 1;
}

void icm_hip_shutdown(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_device = -1;
}

const char *icm_hip_last_error(void) { return g_err.c_str(); }

int icm_hip_release_pooled_memory(void) {
    if (g_device < 0) { set_err("icm_hip_init must be called first"); return -1; }
    if (!HIP_OK(hipSetDevice(g_device))) return -1;
    hipMemPool_t pool = nullptr;
    if (!HIP_OK(hipDeviceGetDefaultMemPool(&pool, g_device))) return -1;
    return HIP_OK(hipMemPoolTrimTo(pool, 0)) ? 0 : -1;
}

IcmHipPlan *icm_hip_plan_create(int n, const double *S, int k, const IcmHipOptions *opts) {
    return plan_create(n, S, k, opts, true);
}

IcmHipPlan *icm_hip_plan_create_ex(int n, const double *S, int k, const IcmHipOptions *opts,
                                   const IcmHipPlanParams *params) {
    return plan_create(n, S, k, opts, true, 256, params);
}

int icm_hip_plan_dry_run(int n, int k, const IcmHipOptions *opts, const IcmHipPlanParams *params,
                         IcmHipPlanSummary *summary, int *retries) {
    if (!params || params->mem_budget_bytes == 0) {
        set_err("dry run needs params->mem_budget_bytes (the simulated free device memory)");
        return -1;
    }
    IcmHipPlan *p = plan_create(n, nullptr, k, opts, false, 256, params, true, params->mem_budget_bytes);
    if (!p) return -1;
    if (summary) icm_hip_plan_summary(p, summary);
    if (retries) *retries = p->retries;
    plan_free(p);
    return 0;
}

void icm_hip_plan_destroy(IcmHipPlan *plan) { plan_free(plan); }

int icm_hip_plan_summary(const IcmHipPlan *p, IcmHipPlanSummary *s) {
    if (!p || !s) { set_err("null argument"); return -1; }
    memset(s, 0, sizeof(*s));
    s->n = p->geo.n; s->k = p->geo.k; s->B = p->geo.B; s->engine = 1; s->n_levels = p->geo.L;
    for (int l = 1; l < p->geo.L; ++l) {
        const int F = p->geo.F[l];
        if (F <= 64) s->n_tier1++;
        else if (F <= icm::kMaxLdsF) s->n_tier2++;
        else s->n_tier3++;
    }
    s->q_batch = p->qb;
    s->planned_peak_vram_bytes = p->mem.bytes;
    return 0;
}

int icm_hip_equity_with_plan(IcmHipPlan *plan, int Q, const double *payout, double *equity,
                             IcmHipRunStats *stats) {
    return equity_host(plan, Q, payout, equity, stats);
}

int icm_hip_equity_with_plan_device(IcmHipPlan *plan, int Q, const double *d_payout, double *d_equity,
                                    IcmHipRunStats *stats) {
    if (!plan || !d_payout || !d_equity) { set_err("null argument"); return -1; }
    if (run_plan(plan, Q, d_payout, stats) != 0) return -1;
    return HIP_OK(hipMemcpy(d_equity, plan->d_equity, sizeof(double) * plan->geo.n, hipMemcpyDeviceToDevice)) ? 0
                                                                                                          : -1;
}

int icm_hip_equity(int n, const double *S, int Q, const double *payout, int k, double *equity) {
    if (Q < 1) { set_err("Q must be positive"); return -1; }
    IcmHipPlan *p = plan_create(n, S, k, nullptr, true, Q);
    if (!p) return -1;
    int rc = equity_host(p, Q, payout, equity, nullptr);
    plan_free(p);
    return rc;
}

int icm_hip_equity_subset(int n, const double *S, int Q, const double *payout, int k, double *equity,
                          const int *targets, int n_targets) {
    if (Q < 1 || !targets || n_targets < 1) { set_err("invalid arguments"); return -1; }
    IcmHipPlan *p = plan_create(n, S, k, nullptr, true, Q);
    if (!p) return -1;
    int rc = set_targets(p, targets, n_targets);
    if (rc == 0) rc = equity_host(p, Q, payout, equity, nullptr);
    plan_free(p);
    return rc;
}

int icm_hip_node_inner(int n, const double *a, const double *payout, int k, double *inner) {
    if (!a || !payout || !inner || n < 1 || k < 1 || k > n) { set_err("invalid arguments"); return -1; }
    std::vector<double> ones(n, 1.0);
    IcmHipPlan *p = plan_create(n, ones.data(), k, nullptr, false, 1);
    if (!p) return -1;
    double *d_a = nullptr, *d_in = nullptr;
    int rc = -1;
    if (dalloc(p, &d_a, n) && dalloc(p, &d_in, n) &&
        HIP_OK(hipMemcpy(d_a, a, sizeof(double) * n, hipMemcpyHostToDevice)) &&
        HIP_OK(hipMemcpy(p->d_payout, payout, sizeof(double) * k, hipMemcpyHostToDevice))) {
        p->d_a_in = d_a;
        p->d_inner = d_in;
        p->raw_node = true;
        p->qb = 1;
        if (run_plan(p, 1, p->d_payout, nullptr) == 0 &&
            HIP_OK(hipMemcpy(inner, d_in, sizeof(double) * n, hipMemcpyDeviceToHost)))
            rc = 0;
    }
    if (d_a) (void)hipFree(d_a);
    p->d_inner = nullptr;
    if (d_in) (void)hipFree(d_in);
    plan_free(p);
    return rc;
}

int icm_hip_vjp_with_plan_device(IcmHipPlan *plan, int Q, const double *d_payout, const double *d_v,
                                 double *d_equity, double *d_dS, IcmHipRunStats *stats) {
    if (!plan || !d_payout || !d_v) { set_err("null argument"); return -1; }
    if (plan->d_active) { set_err("the adjoint pass needs all players (clear the plan's targets)"); return -1; }
    if (run_vjp(plan, Q, d_payout, d_v, stats) != 0) return -1;
    const size_t nb = sizeof(double) * plan->geo.n;
    if (d_equity && !HIP_OK(hipMemcpy(d_equity, plan->d_equity, nb, hipMemcpyDeviceToDevice))) return -1;
    if (d_dS && !HIP_OK(hipMemcpy(d_dS, plan->d_ds, nb, hipMemcpyDeviceToDevice))) return -1;
    return 0;
}

int icm_hip_vjp_with_plan(IcmHipPlan *plan, int Q, const double *payout, const double *v, double *equity,
                          double *dS, IcmHipRunStats *stats) {
    if (!plan || !payout || !v) { set_err("null argument"); return -1; }
    if (plan->d_active) { set_err("the adjoint pass needs all players (clear the plan's targets)"); return -1; }
    if (!HIP_OK(hipSetDevice(g_device))) return -1;
    if (!ensure_vjp(plan)) return -1;
    const size_t nb = sizeof(double) * plan->geo.n;
    // stage v in the dS output buffer (overwritten by the scatter at the end)
    if (!HIP_OK(hipMemcpy(plan->d_payout, payout, sizeof(double) * plan->geo.k, hipMemcpyHostToDevice)) ||
        !HIP_OK(hipMemcpy(plan->d_ds, v, nb, hipMemcpyHostToDevice)))
        return -1;
    if (run_vjp(plan, Q, plan->d_payout, plan->d_ds, stats) != 0) return -1;
    if (equity && !HIP_OK(hipMemcpy(equity, plan->d_equity, nb, hipMemcpyDeviceToHost))) return -1;
    if (dS && !HIP_OK(hipMemcpy(dS, plan->d_ds, nb, hipMemcpyDeviceToHost))) return -1;
    return 0;
}

int icm_hip_equity_vjp(int n, const double *S, int Q, const double *payout, int k, const double *v,
                       double *equity, double *dS) {
    if (Q < 1) { set_err("Q must be positive"); return -1; }
    IcmHipPlan *p = plan_create(n, S, k, nullptr, true, Q);
    if (!p) return -1;
    int rc = icm_hip_vjp_with_plan(p, Q, payout, v, equity, dS, nullptr);
    plan_free(p);
    return rc;
}

int icm_hip_node_vjp(int n, const double *a, const double *c, const double *payout, int k, double *inner,
                     double *H) {
    if (!a || !c || !payout || n < 1 || k < 1 || k > n) { set_err("invalid arguments"); return -1; }
    std::vector<double> ones(n, 1.0);
    IcmHipPlan *p = plan_create(n, ones.data(), k, nullptr, false, 1);
    if (!p) return -1;
    double *d_a = nullptr, *d_c = nullptr, *d_in = nullptr, *d_h = nullptr;
    int rc = -1;
    p->qb = 1;
    if (dalloc(p, &d_a, n) && dalloc(p, &d_c, n) && dalloc(p, &d_in, n) && dalloc(p, &d_h, n) &&
        HIP_OK(hipMemcpy(d_a, a, sizeof(double) * n, hipMemcpyHostToDevice)) &&
        HIP_OK(hipMemcpy(d_c, c, sizeof(double) * n, hipMemcpyHostToDevice)) &&
        HIP_OK(hipMemcpy(p->d_payout, payout, sizeof(double) * k, hipMemcpyHostToDevice))) {
        p->d_a_in = d_a;
        p->d_c_in = d_c;
        p->d_inner = d_in;
        p->d_h = d_h;
        p->raw_node = true;
        if (run_vjp(p, 1, p->d_payout, nullptr, nullptr) == 0 &&
            (!inner || HIP_OK(hipMemcpy(inner, d_in, sizeof(double) * n, hipMemcpyDeviceToHost))) &&
            (!H || HIP_OK(hipMemcpy(H, d_h, sizeof(double) * n, hipMemcpyDeviceToHost))))
            rc = 0;
    }
    p->d_inner = nullptr;
    for (double *x : {d_a, d_c, d_in, d_h}) if (x) (void)hipFree(x);
    plan_free(p);
    return rc;
}

// reference GPU names
int icm_gpu_init(int device_id) { return icm_hip_init(device_id); }
void icm_gpu_shutdown(void) { icm_hip_shutdown(); }
const char *icm_gpu_last_error(void) { return icm_hip_last_error(); }
int icm_gpu_release_pooled_memory(void) { return icm_hip_release_pooled_memory(); }
IcmHipPlan *icm_gpu_plan_create(int n, const double *S, int k, const IcmHipOptions *opts) {
    return icm_hip_plan_create(n, S, k, opts);
}
void icm_gpu_plan_destroy(IcmHipPlan *plan) { icm_hip_plan_destroy(plan); }
int icm_gpu_plan_summary(const IcmHipPlan *plan, IcmHipPlanSummary *s) { return icm_hip_plan_summary(plan, s); }
int icm_gpu_equity_with_plan(IcmHipPlan *plan, int Q, const double *payout, double *equity,
                             IcmHipRunStats *stats) {
    return icm_hip_equity_with_plan(plan, Q, payout, equity, stats);
}
int icm_gpu_equity(int n, const double *S, int Q, const double *payout, int k, double *equity,
                   const IcmHipOptions *opts, IcmHipRunStats *stats) {
    if (Q < 1) { set_err("Q must be positive"); return -1; }
    IcmHipPlan *p = plan_create(n, S, k, opts, true, Q);
    if (!p) return -1;
    int rc = equity_host(p, Q, payout, equity, stats);
    plan_free(p);
    return rc;
}
int icm_gpu_equity_subset(int n, const double *S, int Q, const double *payout, int k, double *equity,
                          const int *targets, int n_targets, const IcmHipOptions *opts, IcmHipRunStats *stats) {
    if (Q < 1 || !targets || n_targets < 1) { set_err("invalid arguments"); return -1; }
    IcmHipPlan *p = plan_create(n, S, k, opts, true, Q);
    if (!p) return -1;
    int rc = set_targets(p, targets, n_targets);
    if (rc == 0) rc = equity_host(p, Q, payout, equity, stats);
    plan_free(p);
    return rc;
}

// CPU-named aliases (python/icm/__init__.py:92-119 binds these three)
void icm_init(const char *wisdom_path) {
    (void)wisdom_path;
    (void)icm_hip_init(0);
}
void icm_equity(int n, const double *S, int Q, const double *payout, int k, double *equity) {
    if (icm_hip_equity(n, S, Q, payout, k, equity) != 0 && equity && n > 0)
        memset(equity, 0, sizeof(double) * n);
}
void icm_equity_subset(int n, const double *S, int Q, const double *payout, int k, double *equity,
                       const int *targets, int n_targets) {
    if (icm_hip_equity_subset(n, S, Q, payout, k, equity, targets, n_targets) != 0 && equity && n > 0)
        memset(equity, 0, sizeof(double) * n);
}

int icm_hip_plan_geometry(int n, int k, int B, int max_levels, int *n_real, int *c, int *G, int *F) {
    Geometry g;
    if (!make_geometry(n, k, B, g)) return -1;
    if (g.L > max_levels) return -1;
    for (int l = 0; l < g.L; ++l) {
        if (n_real) n_real[l] = g.n_real[l];
        if (c) c[l] = g.c[l];
        if (G) G[l] = g.G[l];
        if (F) F[l] = g.F[l];
    }
    return g.L;
}

void icm_hip_make_nodes(int Q, double Smax, double *logv, double *w) { make_nodes(Q, Smax, logv, w); }

int icm_hip_plan_set_targets(IcmHipPlan *plan, const int *targets, int n_targets) {
    if (!plan) { set_err("null plan"); return -1; }
    if (!targets || n_targets <= 0) {   // back to all players
        if (plan->d_active) { (void)hipFree(plan->d_active); plan->d_active = nullptr; }
        if (plan->d_lists) { (void)hipFree(plan->d_lists); plan->d_lists = nullptr; }
        return 0;
    }
    return set_targets(plan, targets, n_targets);
}

int icm_hip_plan_set_schoolbook_max(IcmHipPlan *plan, int cmax) {
    if (!plan) { set_err("null plan"); return -1; }
    plan->sb_max = cmax < 0 ? 0 : cmax;
    return 0;
}

int icm_hip_plan_get_schoolbook_max(const IcmHipPlan *plan) { return plan ? plan->sb_max : -1; }

int icm_hip_plan_set_small_path(IcmHipPlan *plan, int mode) {
    if (!plan) { set_err("null plan"); return -1; }
    if (mode < -1 || mode > 1) { set_err("small-path mode must be -1 (tuned), 0 (off) or 1 (on)"); return -1; }
    plan->small_mode = mode;
    return 0;
}

int icm_hip_plan_set_column_fusion(IcmHipPlan *plan, int on) {
    if (!plan) { set_err("null plan"); return -1; }
    plan->fuse_columns = on < 0 ? 15 : (on & 15);
    return 0;
}

long long icm_hip_plan_poisoned_bytes(const IcmHipPlan *plan) { return plan ? plan->poisoned_bytes : -1; }

int icm_hip_selfcheck_row_layouts(void) {
    std::string why;
    if (icm::selfcheck_row_layouts(why)) return 0;
    set_err("row layout self-check: " + why);
    return -1;
}

int icm_hip_plan_set_adjoint_checkpoint(IcmHipPlan *plan, int interval) {
    if (!plan) { set_err("null plan"); return -1; }
    if (!(interval == -1 || interval == 0 || (interval >= 2 && interval <= 64))) {
        set_err("adjoint checkpoint interval must be -1 (auto), 0 (keep every level) or 2 .. 64");
        return -1;
    }
    if (interval != plan->ckpt_mode && plan->qv > 0) {
        if (!HIP_OK(hipSetDevice(g_device)) || !HIP_OK(hipStreamSynchronize(plan->stream))) return -1;
        free_vjp(plan, plan->vjp_base);   // reallocated on the next adjoint call
    }
    plan->ckpt_mode = interval;
    return 0;
}

int icm_hip_plan_adjoint_layout(const IcmHipPlan *plan, int *interval, int *q_planes) {
    if (!plan) { set_err("null plan"); return -1; }
    if (plan->qv <= 0) { set_err("no adjoint buffers yet (they are sized on the first adjoint call)"); return -1; }
    if (interval) *interval = plan->ckpt;
    if (q_planes) *q_planes = std::min(plan->qb, plan->qv);
    return 0;
}

int icm_hip_plan_set_poison(IcmHipPlan *plan, int on) {
    if (!plan) { set_err("null plan"); return -1; }
    plan->poison = on != 0;
    return 0;
}

int icm_hip_plan_get_small_path(const IcmHipPlan *plan) {
    if (!plan) return -1;
    return use_small(plan) ? 1 : 0;
}

int icm_hip_plan_set_spectrum_cache(IcmHipPlan *plan, int on) {
    if (!plan) { set_err("null plan"); return -1; }
    if (!on) {
        spec_release(plan);
        return 0;
    }
    if (plan->d_spec || plan->spec_elems == 0) return 0;
    if (!spec_alloc(plan)) {
        set_err("not enough device memory for the spectrum cache at this plan's q-batch");
        return -1;
    }
    return 0;
}

int icm_hip_plan_get_spectrum_cache(const IcmHipPlan *plan) { return plan ? (plan->d_spec ? 1 : 0) : -1; }

int icm_hip_plan_set_profiling(IcmHipPlan *plan, int on) {
    if (!plan) { set_err("null plan"); return -1; }
    plan->prof.enabled = on != 0;
    plan->prof.clear();
    return 0;
}

int icm_hip_plan_launch_count(const IcmHipPlan *plan) { return plan ? (int)plan->prof.log.size() : -1; }

int icm_hip_plan_launch_stat(const IcmHipPlan *plan, int index, char *name, int name_len, int *level, double *ns,
                             double *alg_bytes) {
    if (!plan || index < 0 || index >= (int)plan->prof.log.size()) { set_err("bad launch index"); return -1; }
    const auto &a = plan->prof.log[index];
    if (name && name_len > 0) {
        strncpy(name, a.name.c_str(), name_len - 1);
        name[name_len - 1] = 0;
    }
    if (level) *level = a.tag;
    if (ns) *ns = a.ns;
    if (alg_bytes) *alg_bytes = a.bytes;
    return 0;
}

int icm_hip_plan_kernel_count(const IcmHipPlan *plan) {
    return plan ? (int)plan->prof.agg.size() : -1;
}

int icm_hip_plan_kernel_stat(const IcmHipPlan *plan, int index, char *name, int name_len, int *launches,
                             double *total_ns, double *alg_bytes, double *alg_flops) {
    if (!plan || index < 0 || index >= (int)plan->prof.agg.size()) { set_err("bad kernel index"); return -1; }
    const auto &a = plan->prof.agg[index];
    if (name && name_len > 0) {
        strncpy(name, a.name.c_str(), name_len - 1);
        name[name_len - 1] = 0;
    }
    if (launches) *launches = a.launches;
    if (total_ns) *total_ns = a.ns;
    if (alg_bytes) *alg_bytes = a.bytes;
    if (alg_flops) *alg_flops = a.flops;
    return 0;
}

}  // extern "C"
