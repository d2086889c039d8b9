// icm_batch.cpp -- ragged batches: many independent ICM instances of uneven
// n and k solved together.
//
// Every instance keeps its own tree (geometry of icm_host.cpp with one block
// size B for the whole batch), but the work of all instances runs in one
// launch per phase: leaf build, one launch per tree level for the build
// (levels 1 .. Lmax-2) and for the propagate (Lmax-1 .. 1), leaf extract,
// q-reduction and scatter.  A level launch is a flat list of work items
// (instance job, first tile item); each 4096-point LDS tile is filled with
// (q-plane, parent pair) items of one instance, so a small instance's level
// costs a few tiles for all Q planes instead of one mostly empty tile per
// plane.  The first run with a given Q records the launches as one HIP graph;
// later runs replay it (one hipGraphLaunch per batch).
//
// With block size 64 (the tuned default) a q-pass can instead run as ONE
// persistent launch (k_batch_persist): workgroups take the same work items,
// plus leaf, extract and reduction items, by ticket, and each instance moves
// to its next level as soon as its own previous level is done (per-instance
// completion counters), with no barrier between levels.  The g arrays then
// give every instance its own region per level parity, so instances at
// different levels never share memory.
//
// Instances with a level above one LDS transform (F > 4096: k > 2048 and
// n > 4096) cannot share those launches; a batch that holds some becomes a
// composite: the other instances run as an inner batch, and each large one
// through a single-instance plan (four-step levels), writing its segment of
// the batch's flat arrays.
//
// Reference semantics per instance are those of icm_equity
// (src/cpu/icm.h:26-41, src/cpu/icm.c:2523-2695); the batch entry itself has
// no counterpart in the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/icm_hip.h"
#include "icm_common.hpp"
#include "icm_kernels.hpp"

using namespace icmh;

struct IcmHipBatchPlan {
    int ninst = 0, Ntot = 0, Ktot = 0, NB = 0, Lmax = 0, B = 32;
    std::vector<int> n_off, k_off;
    std::vector<Geometry> geo;
    std::vector<int> perm;               // global sorted position -> flat original index
    std::vector<double> S_sorted;        // per instance, descending, concatenated
    std::vector<double> Smax;
    std::vector<int> boff;               // first global block of each instance
    std::vector<int> blk_inst;           // global block -> instance
    int qb = 0;                          // q-planes per pass
    int nchunk = 1;
    int twShift = 12;
    int sb_max = 0;                      // schoolbook levels: c_{l-1} <= sb_max
    bool use_graph = true;
    bool verbose = false;
    bool persistent = false;             // one persistent launch per q-pass (needs persist_ok)
    bool persist_ok = false;             // block size the persistent kernel is built for
    int persist_grid = 0;                // resident workgroups of the persistent kernel
    int spin_limit = 1 << 20;            // bound of a persistent item's dependency wait (polls, < 1 s)
    int *d_ctr = nullptr;                // [ninst] finished persistent items per instance
    int *d_sched = nullptr;              // [kPSchedSize] queue tickets, workgroups done, error flag
    // arena offsets (doubles) per level and instance, sized for qb planes
    std::vector<std::vector<long long>> coff, goff;
    std::vector<long long> coef_elems;   // per level
    long long g_elems[2] = {0, 0};
    // device memory
    double2 *d_tw = nullptr;
    double *d_S = nullptr;
    int *d_perm = nullptr;
    double *d_payout = nullptr;
    double *d_nodes = nullptr;           // [ninst][2][Q]
    int nodes_Q = 0;
    int nodes_alloc_Q = 0;               // Q that d_nodes was sized for
    icm::BatchInst *d_inst = nullptr;
    int *d_blk_inst = nullptr;
    double *d_blocks = nullptr;          // [qb][NB][B+1]
    std::vector<double *> d_coef;        // level 1 .. Lmax-2
    double *d_g[2] = {nullptr, nullptr};
    double *d_partial = nullptr;         // [nchunk][Ntot]
    double *d_acc = nullptr;             // [Ntot] sorted
    double *d_equity = nullptr;          // [Ntot] original flat order
    // launch schedule per q-pass size
    struct Pass {
        int nq = 0;
        icm::LevelJob *d_jobs = nullptr;
        int2 *d_items = nullptr;
        size_t dev_bytes = 0;
        // per level l: build items [b_off[l], b_off[l] + b_cnt[l]), prop items
        // likewise; transform levels, then schoolbook levels (bs_*, ps_*)
        std::vector<int> b_off, b_cnt, p_off, p_cnt;
        std::vector<double> b_bytes, b_flops, p_bytes, p_flops;
        std::vector<int> bs_off, bs_cnt, ps_off, ps_cnt;
        std::vector<double> bs_bytes, bs_flops, ps_bytes, ps_flops;
        // persistent launch: ordered items and per-instance need counts
        icm::PItem *d_pitems = nullptr;
        int n_pitems = 0;
        int *d_need = nullptr;           // [2][ninst]: before extract, before finalize
        int *d_qtab = nullptr;           // queue item offsets [kPMaxQueues + 1], block / player ends [kPMaxQueues][2]
        std::vector<int> qcut;           // first instance of each queue (+ ninst)
        int nch = 1, qpc = 1;            // q-chunks of the extract and q-planes per chunk
    };
    std::vector<Pass> passes;            // index 0: nq = qb, 1: the remainder (if any)
    // adjoint pass (allocated on the first VJP call)
    bool vjp_ready = false;
    double *d_dblocks = nullptr;         // [qb][NB][B+1] tangent block products
    std::vector<double *> d_dcoef;       // tangent products per level (same offsets as d_coef)
    std::vector<std::vector<long long>> bvoff;
    long long bv_elems[2] = {0, 0};
    double *d_bv[2] = {nullptr, nullptr};// B adjoint ping-pong by level parity
    double *d_vin = nullptr;             // cotangent, flat original order
    double *d_v = nullptr;               // cotangent, sorted
    double *d_partial_ds = nullptr, *d_acc_ds = nullptr, *d_ds = nullptr;
    std::vector<Pass> vpasses;
    hipGraph_t graph[2] = {nullptr, nullptr};       // [0] forward, [1] adjoint
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int graph_Q[2] = {0, 0};
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DevMem mem;                          // device bytes held (cap accounting)
    int retries = 0;                     // q-pass halvings at creation (out-of-memory policy)
    KernelProf prof;
    // composite batch: instances whose levels fit one LDS transform run as
    // `inner` (flat index maps into this batch's arrays), the others through
    // single-instance plans; d_payout / d_equity / d_vin / d_ds hold the
    // whole batch in its flat original order
    bool composite = false;
    IcmHipBatchPlan *inner = nullptr;
    std::vector<int> large_ids;
    std::vector<IcmHipPlan *> large;
    int *d_kmap = nullptr;               // inner payout index -> batch payout index
    int *d_nmap = nullptr;               // inner player index -> batch player index
};

namespace {

template <class T>
bool balloc(IcmHipBatchPlan *p, T **ptr, size_t count) {
    return p->mem.alloc(ptr, count);
}

void drop_graph(IcmHipBatchPlan *p, int mode) {
    if (p->gexec[mode]) (void)hipGraphExecDestroy(p->gexec[mode]);
    if (p->graph[mode]) (void)hipGraphDestroy(p->graph[mode]);
    p->gexec[mode] = nullptr;
    p->graph[mode] = nullptr;
    p->graph_Q[mode] = 0;
}

void drop_graphs(IcmHipBatchPlan *p) {
    drop_graph(p, 0);
    drop_graph(p, 1);
}

void free_passes(IcmHipBatchPlan *p, std::vector<IcmHipBatchPlan::Pass> &v) {
    for (auto &ps : v) {
        if (ps.d_jobs) (void)hipFree(ps.d_jobs);
        if (ps.d_items) (void)hipFree(ps.d_items);
        if (ps.d_pitems) (void)hipFree(ps.d_pitems);
        if (ps.d_need) (void)hipFree(ps.d_need);
        if (ps.d_qtab) (void)hipFree(ps.d_qtab);
        p->mem.bytes -= std::min(p->mem.bytes, ps.dev_bytes);
    }
    v.clear();
}

// Frees the plan's device buffers (forward, adjoint, node and job tables)
// and clears the byte count; graphs, stream and events stay.
void free_buffers(IcmHipBatchPlan *p) {
    drop_graphs(p);
    free_passes(p, p->passes);
    free_passes(p, p->vpasses);
    auto f = [](auto *&x) { if (x) (void)hipFree((void *)x); x = nullptr; };
    f(p->d_tw); f(p->d_S); f(p->d_perm); f(p->d_payout); f(p->d_nodes); f(p->d_inst); f(p->d_blk_inst);
    f(p->d_blocks); for (double *&c : p->d_coef) f(c);
    f(p->d_g[0]); f(p->d_g[1]); f(p->d_partial); f(p->d_acc); f(p->d_equity);
    f(p->d_dblocks); for (double *&c : p->d_dcoef) f(c);
    f(p->d_bv[0]); f(p->d_bv[1]); f(p->d_vin); f(p->d_v); f(p->d_partial_ds); f(p->d_acc_ds); f(p->d_ds);
    f(p->d_ctr); f(p->d_sched);
    p->nodes_Q = 0;
    p->vjp_ready = false;
    p->mem.bytes = 0;
}

void batch_free(IcmHipBatchPlan *p) {
    if (!p) return;
    if (p->inner) batch_free(p->inner);
    for (IcmHipPlan *q : p->large) icm_hip_plan_destroy(q);
    auto f = [](auto *&x) { if (x) (void)hipFree((void *)x); x = nullptr; };
    f(p->d_kmap); f(p->d_nmap);
    free_buffers(p);
    if (p->ev0) (void)hipEventDestroy(p->ev0);
    if (p->ev1) (void)hipEventDestroy(p->ev1);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

// Level-l coefficients of instance i (l = 0: its blocks in the global array).
icm::CoefView coef_view(const IcmHipBatchPlan *p, int i, int l) {
    const Geometry &g = p->geo[i];
    icm::CoefView v{};
    if (l == 0) {
        v.base = p->d_blocks + (long long)p->boff[i] * (p->B + 1);
        v.plane = (long long)p->NB * (p->B + 1);
        v.stride = p->B + 1;
        v.count = g.c[0];
    } else {
        v.base = p->d_coef[l] + p->coff[l][i];
        v.plane = (long long)g.n_real[l] * g.c[l];
        v.stride = g.c[l];
        v.count = g.c[l];
    }
    return v;
}

icm::CoefView g_view(const IcmHipBatchPlan *p, int i, int l) {
    const Geometry &g = p->geo[i];
    if (l == g.L - 1)   // root: the payout, shared by every q-plane
        return icm::CoefView{p->d_payout + p->k_off[i], 0, g.k, g.k};
    return icm::CoefView{p->d_g[l & 1] + p->goff[l][i], (long long)g.n_real[l] * g.G[l], g.G[l], g.G[l]};
}

// Tangent products of instance i at level l (l = 0: its tangent blocks).
icm::CoefView tcoef_view(const IcmHipBatchPlan *p, int i, int l) {
    icm::CoefView v = coef_view(p, i, l);
    v.base = (l == 0) ? p->d_dblocks + (long long)p->boff[i] * (p->B + 1) : p->d_dcoef[l] + p->coff[l][i];
    return v;
}

// B adjoint of instance i at level l (zero at the root: base null).
icm::CoefView bv_view(const IcmHipBatchPlan *p, int i, int l) {
    const Geometry &g = p->geo[i];
    if (l >= g.L - 1) return icm::CoefView{};
    return icm::CoefView{p->d_bv[l & 1] + p->bvoff[l][i], (long long)g.n_real[l] * g.c[l], g.c[l], g.c[l]};
}

double fft5(int F) { return 5.0 * F * (double)ilog2i(F); }

// Adjoint job tables: fused levels, one parent per tile item.
bool make_vjp_pass(IcmHipBatchPlan *p, int nq, IcmHipBatchPlan::Pass &ps) {
    ps.nq = nq;
    const int Lm = p->Lmax;
    std::vector<icm::LevelJob> jobs;
    std::vector<int2> items;
    ps.b_off.assign(Lm, 0); ps.b_cnt.assign(Lm, 0); ps.p_off.assign(Lm, 0); ps.p_cnt.assign(Lm, 0);
    ps.b_bytes.assign(Lm, 0); ps.b_flops.assign(Lm, 0); ps.p_bytes.assign(Lm, 0); ps.p_flops.assign(Lm, 0);
    auto add_items = [&](int job, int nparent, int F) {
        const long long nv = (long long)nq * nparent;
        const int per = icm::kMaxLdsF / F;
        for (long long v0 = 0; v0 < nv; v0 += per) items.push_back(make_int2(job, (int)v0));
    };
    for (int l = 1; l <= Lm - 2; ++l) {
        ps.b_off[l] = (int)items.size();
        for (int i = 0; i < p->ninst; ++i) {
            const Geometry &g = p->geo[i];
            if (l > g.L - 2) continue;
            icm::LevelJob J{};
            J.child = coef_view(p, i, l - 1);
            J.tchild = tcoef_view(p, i, l - 1);
            const long long pl = (long long)g.n_real[l] * g.c[l];
            J.parent = icm::CoefOut{p->d_coef[l] + p->coff[l][i], pl, g.c[l], g.c[l]};
            J.tparent = icm::CoefOut{p->d_dcoef[l] + p->coff[l][i], pl, g.c[l], g.c[l]};
            J.nr_child = g.n_real[l - 1];
            J.nr_parent = g.n_real[l];
            J.np = g.n_real[l];
            J.F = g.F[l];
            J.logF = ilog2i(g.F[l]);
            J.nq = nq;
            add_items((int)jobs.size(), g.n_real[l], J.F);
            jobs.push_back(J);
            ps.b_bytes[l] += 16.0 * nq * ((double)g.n_real[l - 1] * g.c[l - 1] + (double)g.n_real[l] * g.c[l]);
            ps.b_flops[l] += (double)nq * g.n_real[l] * 3.0 * fft5(J.F);
        }
        ps.b_cnt[l] = (int)items.size() - ps.b_off[l];
    }
    for (int l = Lm - 1; l >= 1; --l) {
        ps.p_off[l] = (int)items.size();
        for (int i = 0; i < p->ninst; ++i) {
            const Geometry &g = p->geo[i];
            if (l > g.L - 1) continue;
            icm::LevelJob J{};
            J.child = coef_view(p, i, l - 1);
            J.tchild = tcoef_view(p, i, l - 1);
            J.gpar = g_view(p, i, l);
            J.bpar = bv_view(p, i, l);
            const int cc = g.c[l - 1];
            J.gchild = icm::CoefOut{p->d_g[(l - 1) & 1] + p->goff[l - 1][i],
                                    (long long)g.n_real[l - 1] * g.G[l - 1], g.G[l - 1], g.G[l - 1]};
            J.bchild = icm::CoefOut{p->d_bv[(l - 1) & 1] + p->bvoff[l - 1][i], (long long)g.n_real[l - 1] * cc, cc,
                                    cc};
            J.gfix = (g.F[l] < 2 * cc - 1) ? 1 : 0;
            J.nr_child = g.n_real[l - 1];
            J.nr_parent = g.n_real[l];
            J.np = g.n_real[l];
            J.F = g.F[l];
            J.logF = ilog2i(g.F[l]);
            J.nq = nq;
            add_items((int)jobs.size(), g.n_real[l], J.F);
            jobs.push_back(J);
            const double ain = (l == g.L - 1) ? 8.0 * g.k : 8.0 * nq * (double)g.n_real[l] * (g.G[l] + g.c[l]);
            ps.p_bytes[l] += ain + 8.0 * nq * (double)g.n_real[l - 1] * (3.0 * cc + g.G[l - 1]);
            ps.p_flops[l] += (double)nq * g.n_real[l] * 5.0 * fft5(J.F);
        }
        ps.p_cnt[l] = (int)items.size() - ps.p_off[l];
    }
    if (items.size() > 0x7fffffffULL) { set_err("batch too large: work items exceed 2^31"); return false; }
    const size_t before = p->mem.bytes;
    if (!balloc(p, &ps.d_jobs, jobs.size()) || !balloc(p, &ps.d_items, items.size())) return false;
    ps.dev_bytes = p->mem.bytes - before;
    return (jobs.empty() || HIP_OK(hipMemcpy(ps.d_jobs, jobs.data(), sizeof(icm::LevelJob) * jobs.size(),
                                             hipMemcpyHostToDevice))) &&
           (items.empty() || HIP_OK(hipMemcpy(ps.d_items, items.data(), sizeof(int2) * items.size(),
                                              hipMemcpyHostToDevice)));
}

// q-chunks of a pass's leaf extract: (chunks, q-planes per chunk).
void extract_chunks(const IcmHipBatchPlan *p, int nq, int &nch, int &qpc) {
    const int nchunk = std::min(p->nchunk, nq);
    qpc = (nq + nchunk - 1) / nchunk;
    nch = (nq + qpc - 1) / qpc;
}

// The persistent launch's item lists for one pass, one per XCD queue.  The
// instances are cut into kPMaxQueues contiguous ranges of about equal work
// (level items, schoolbook ones at half weight, and blocks); queue x holds,
// for its instances only: leaf items over its blocks, every level's work
// items (in launch order: build levels up, propagate levels down), extract
// items over its blocks and finalize items over its players.  An instance's
// items of one phase all need its finished items of the earlier phases;
// those come first in its queue, so every wait is on smaller tickets of the
// same queue, and no item reads what another XCD writes.
bool make_persist_items(IcmHipBatchPlan *p, IcmHipBatchPlan::Pass &ps, const std::vector<int2> &items,
                        const std::vector<int> &job_inst) {
    const int ni = p->ninst, NB = p->NB, Lm = p->Lmax, NQ = icm::kPMaxQueues;
    // level steps in launch order: (offset, count, kind)
    struct Rng { int off, cnt, kind; };
    std::vector<std::vector<Rng>> steps;
    for (int l = 1; l <= Lm - 2; ++l)
        steps.push_back({{ps.b_off[l], ps.b_cnt[l], icm::kPBuildLds}, {ps.bs_off[l], ps.bs_cnt[l], icm::kPBuildSb}});
    for (int l = Lm - 1; l >= 1; --l)
        steps.push_back({{ps.p_off[l], ps.p_cnt[l], icm::kPPropLds}, {ps.ps_off[l], ps.ps_cnt[l], icm::kPPropSb}});
    // work per instance and the queue cuts
    std::vector<double> wi(ni, 0.0);
    double witems = 0.0;
    for (const auto &st : steps)
        for (const Rng &r : st)
            for (int x = r.off; x < r.off + r.cnt; ++x) {
                const double w = (r.kind == icm::kPBuildSb || r.kind == icm::kPPropSb) ? 0.5 : 1.0;
                wi[job_inst[items[x].x]] += w;
                witems += w;
            }
    std::vector<double> cost(ni);
    double total = 0.0;
    for (int i = 0; i < ni; ++i) {
        cost[i] = (witems > 0 ? 0.7 * wi[i] / witems : 0.0) + (witems > 0 ? 0.3 : 1.0) * p->geo[i].nb / NB;
        total += cost[i];
    }
    std::vector<int> cut(NQ + 1, ni);
    cut[0] = 0;
    {
        double acc = 0.0;
        int x = 1;
        for (int i = 0; i < ni && x < NQ; ++i) {
            acc += cost[i];
            while (x < NQ && acc >= total * x / NQ) cut[x++] = i + 1;
        }
        for (int q = 1; q <= NQ; ++q) cut[q] = std::max(cut[q], cut[q - 1]);
    }
    std::vector<icm::PItem> pit;
    std::vector<int> base(ni, 0), step(ni, 0), touched, need(2 * (size_t)ni, 0), qtab(3 * NQ + 1, 0);
    extract_chunks(p, ps.nq, ps.nch, ps.qpc);
    const int exb = icm::persist_extract_blocks(p->B);
    auto inst_of_player = [&](int j) {
        return (int)(std::upper_bound(p->n_off.begin(), p->n_off.end(), j) - p->n_off.begin()) - 1;
    };
    for (int q = 0; q < NQ; ++q) {
        qtab[q] = (int)pit.size();
        const int ilo = cut[q], ihi = cut[q + 1];
        const int blo = ilo < ni ? p->boff[ilo] : NB, bhi = ihi < ni ? p->boff[ihi] : NB;
        const int plo = p->n_off[ilo], phi = p->n_off[ihi];
        qtab[NQ + 1 + 2 * q] = bhi;
        qtab[NQ + 1 + 2 * q + 1] = phi;
        if (ilo >= ihi) continue;
        for (int gb0 = blo; gb0 < bhi; gb0 += icm::kPLeafBlocks) {
            const int i0 = p->blk_inst[gb0], i1 = p->blk_inst[std::min(gb0 + icm::kPLeafBlocks, bhi) - 1];
            pit.push_back(icm::PItem{icm::kPLeaf | (i0 << 3), gb0, 0, i1});
            for (int i = i0; i <= i1; ++i) ++base[i];
        }
        for (const auto &st : steps) {
            for (const Rng &r : st)
                for (int x = r.off; x < r.off + r.cnt; ++x) {
                    const int i = job_inst[items[x].x];
                    if (i < ilo || i >= ihi) continue;
                    pit.push_back(icm::PItem{r.kind | (i << 3), items[x].x, items[x].y, base[i]});
                    if (step[i]++ == 0) touched.push_back(i);
                }
            for (int i : touched) { base[i] += step[i]; step[i] = 0; }
            touched.clear();
        }
        for (int i = ilo; i < ihi; ++i) need[i] = base[i];
        for (int c = 0; c < ps.nch; ++c)
            for (int gb0 = blo; gb0 < bhi; gb0 += exb) {
                const int i0 = p->blk_inst[gb0], i1 = p->blk_inst[std::min(gb0 + exb, bhi) - 1];
                pit.push_back(icm::PItem{icm::kPExtract | (i0 << 3), gb0, c, i1});
                for (int i = i0; i <= i1; ++i) ++base[i];
            }
        for (int i = ilo; i < ihi; ++i) need[ni + i] = base[i];
        for (int j0 = plo; j0 < phi; j0 += icm::kPFinalPlayers) {
            const int i0 = inst_of_player(j0), i1 = inst_of_player(std::min(j0 + icm::kPFinalPlayers, phi) - 1);
            pit.push_back(icm::PItem{icm::kPFinal | (i0 << 3), j0, 0, i1});
        }
    }
    qtab[NQ] = (int)pit.size();
    for (const icm::PItem &x : pit) {   // a multi-instance item's waits and counts are one per thread
        const int k = x.kind_inst & 7;
        if ((k == icm::kPLeaf || k == icm::kPExtract || k == icm::kPFinal) && x.need - (x.kind_inst >> 3) >= 512) {
            set_err("persistent batch: an item spans more than 512 instances");
            return false;
        }
    }
    ps.n_pitems = (int)pit.size();
    ps.qcut = cut;
    const size_t before = p->mem.bytes;
    if (!balloc(p, &ps.d_pitems, std::max<size_t>(1, pit.size())) || !balloc(p, &ps.d_need, need.size()) ||
        !balloc(p, &ps.d_qtab, qtab.size()))
        return false;
    ps.dev_bytes += p->mem.bytes - before;
    return (pit.empty() ||
            HIP_OK(hipMemcpy(ps.d_pitems, pit.data(), sizeof(icm::PItem) * pit.size(), hipMemcpyHostToDevice))) &&
           HIP_OK(hipMemcpy(ps.d_need, need.data(), sizeof(int) * need.size(), hipMemcpyHostToDevice)) &&
           HIP_OK(hipMemcpy(ps.d_qtab, qtab.data(), sizeof(int) * qtab.size(), hipMemcpyHostToDevice));
}

// Job tables and work items for passes of nq q-planes.
bool make_pass(IcmHipBatchPlan *p, int nq, IcmHipBatchPlan::Pass &ps) {
    ps.nq = nq;
    const int Lm = p->Lmax;
    std::vector<icm::LevelJob> jobs;
    std::vector<int> job_inst;
    std::vector<int2> items;
    for (auto *v : {&ps.b_off, &ps.b_cnt, &ps.p_off, &ps.p_cnt, &ps.bs_off, &ps.bs_cnt, &ps.ps_off, &ps.ps_cnt})
        v->assign(Lm, 0);
    for (auto *v : {&ps.b_bytes, &ps.b_flops, &ps.p_bytes, &ps.p_flops, &ps.bs_bytes, &ps.bs_flops, &ps.ps_bytes,
                    &ps.ps_flops})
        v->assign(Lm, 0.0);
    // work items of one job: tile items over its (q-plane, unit) space, `per` units each
    auto add_items = [&](std::vector<int2> &dst, int job, int units, int per) {
        const long long nv = (long long)nq * units;
        for (long long v0 = 0; v0 < nv; v0 += per) dst.push_back(make_int2(job, (int)v0));
    };
    auto sb_ok = [&](const Geometry &g, int l) {
        const int cc = g.c[l - 1];
        return cc <= p->sb_max && cc <= icm::kSbMaxC && icm::sb_stride_prop(cc, g.G[l - 1]) <= icm::kSbLds;
    };
    for (int l = 1; l <= Lm - 2; ++l) {
        std::vector<int2> fft, sb;
        for (int i = 0; i < p->ninst; ++i) {
            const Geometry &g = p->geo[i];
            if (l > g.L - 2) continue;
            icm::LevelJob J{};
            J.child = coef_view(p, i, l - 1);
            J.parent = icm::CoefOut{p->d_coef[l] + p->coff[l][i], (long long)g.n_real[l] * g.c[l], g.c[l], g.c[l]};
            J.nr_child = g.n_real[l - 1];
            J.nr_parent = g.n_real[l];
            J.np = g.n_real[l];
            J.F = g.F[l];
            J.logF = ilog2i(g.F[l]);
            J.nq = nq;
            const double by = 8.0 * nq * ((double)g.n_real[l - 1] * g.c[l - 1] + (double)g.n_real[l] * g.c[l]);
            if (sb_ok(g, l)) {
                add_items(sb, (int)jobs.size(), g.n_real[l], icm::sb_ppw(icm::sb_stride_build(g.c[l - 1])));
                ps.bs_bytes[l] += by;
                ps.bs_flops[l] += 2.0 * nq * (double)g.n_real[l] * g.c[l - 1] * g.c[l];
            } else {
                const int npair = (g.n_real[l] + 1) / 2;
                add_items(fft, (int)jobs.size(), npair, icm::kMaxLdsF / J.F);
                ps.b_bytes[l] += by;
                ps.b_flops[l] += (double)nq * npair * 3.0 * fft5(J.F);
            }
            jobs.push_back(J);
            job_inst.push_back(i);
        }
        ps.b_off[l] = (int)items.size();
        ps.b_cnt[l] = (int)fft.size();
        items.insert(items.end(), fft.begin(), fft.end());
        ps.bs_off[l] = (int)items.size();
        ps.bs_cnt[l] = (int)sb.size();
        items.insert(items.end(), sb.begin(), sb.end());
    }
    for (int l = Lm - 1; l >= 1; --l) {
        std::vector<int2> fft, sb;
        for (int i = 0; i < p->ninst; ++i) {
            const Geometry &g = p->geo[i];
            if (l > g.L - 1) continue;
            icm::LevelJob J{};
            J.child = coef_view(p, i, l - 1);
            J.gpar = g_view(p, i, l);
            J.gchild = icm::CoefOut{p->d_g[(l - 1) & 1] + p->goff[l - 1][i],
                                    (long long)g.n_real[l - 1] * g.G[l - 1], g.G[l - 1], g.G[l - 1]};
            J.nr_child = g.n_real[l - 1];
            J.nr_parent = g.n_real[l];
            J.np = g.n_real[l];
            J.F = g.F[l];
            J.logF = ilog2i(g.F[l]);
            J.nq = nq;
            const double gin = (l == g.L - 1) ? 8.0 * g.k : 8.0 * nq * (double)g.n_real[l] * g.G[l];
            const double by = gin + 8.0 * nq * (double)g.n_real[l - 1] * (g.c[l - 1] + g.G[l - 1]);
            if (sb_ok(g, l)) {
                add_items(sb, (int)jobs.size(), g.n_real[l],
                          icm::sb_ppw(icm::sb_stride_prop(g.c[l - 1], g.G[l - 1])));
                ps.ps_bytes[l] += by;
                ps.ps_flops[l] += 4.0 * nq * (double)g.n_real[l] * g.c[l - 1] * g.G[l - 1];
            } else {
                const int npair = (g.n_real[l] + 1) / 2;
                add_items(fft, (int)jobs.size(), npair, icm::kMaxLdsF / J.F);
                ps.p_bytes[l] += by;
                ps.p_flops[l] += (double)nq * npair * 5.0 * fft5(J.F);
            }
            jobs.push_back(J);
            job_inst.push_back(i);
        }
        ps.p_off[l] = (int)items.size();
        ps.p_cnt[l] = (int)fft.size();
        items.insert(items.end(), fft.begin(), fft.end());
        ps.ps_off[l] = (int)items.size();
        ps.ps_cnt[l] = (int)sb.size();
        items.insert(items.end(), sb.begin(), sb.end());
    }
    if (items.size() > 0x7fffffffULL) { set_err("batch too large: work items exceed 2^31"); return false; }
    const size_t before = p->mem.bytes;
    if (!balloc(p, &ps.d_jobs, jobs.size()) || !balloc(p, &ps.d_items, items.size())) return false;
    ps.dev_bytes = p->mem.bytes - before;
    return (jobs.empty() || HIP_OK(hipMemcpy(ps.d_jobs, jobs.data(), sizeof(icm::LevelJob) * jobs.size(),
                                             hipMemcpyHostToDevice))) &&
           (items.empty() || HIP_OK(hipMemcpy(ps.d_items, items.data(), sizeof(int2) * items.size(),
                                              hipMemcpyHostToDevice))) &&
           (!p->persist_ok || make_persist_items(p, ps, items, job_inst));
}

IcmHipBatchPlan *batch_create(int n_inst, const int *n_off, const double *S, const int *k_off,
                              const IcmHipOptions *opts, const IcmHipPlanParams *params = nullptr);

// A batch with instances whose levels exceed one LDS transform: those get
// single-instance plans (each planned for an equal share of the memory
// fraction), the rest an inner batch over compacted offsets.
IcmHipBatchPlan *composite_create(int n_inst, const int *n_off, const double *S, const int *k_off,
                                  const IcmHipOptions *opts, const IcmHipPlanParams *params, const PlanCfg &cfg,
                                  const std::vector<int> &large_ids) {
    IcmHipBatchPlan *p = new IcmHipBatchPlan();
    p->composite = true;
    p->ninst = n_inst;
    p->Ntot = n_off[n_inst];
    p->Ktot = k_off[n_inst];
    p->n_off.assign(n_off, n_off + n_inst + 1);
    p->k_off.assign(k_off, k_off + n_inst + 1);
    p->large_ids = large_ids;
    p->verbose = cfg.verbose;
    const int nl = (int)large_ids.size();
    IcmHipPlanParams prm{};
    if (params) prm = *params;
    prm.mem_fraction = cfg.frac / (nl + 1);
    if (cfg.cap) prm.mem_budget_bytes = cfg.cap / (nl + 1);
    std::vector<char> is_large(n_inst, 0);
    for (int i : large_ids) {
        is_large[i] = 1;
        IcmHipPlan *q = icm_hip_plan_create_ex(n_off[i + 1] - n_off[i], S + n_off[i], k_off[i + 1] - k_off[i], opts,
                                               &prm);
        if (!q) {
            const std::string why = g_err;
            char b[96];
            snprintf(b, sizeof b, "batch instance %d (n=%d, k=%d): ", i, n_off[i + 1] - n_off[i],
                     k_off[i + 1] - k_off[i]);
            set_err(b + why);
            batch_free(p);
            return nullptr;
        }
        p->large.push_back(q);
    }
    // the other instances, compacted
    std::vector<int> no{0}, ko{0}, nmap, kmap;
    std::vector<double> Sc;
    for (int i = 0; i < n_inst; ++i) {
        if (is_large[i]) continue;
        for (int j = n_off[i]; j < n_off[i + 1]; ++j) { Sc.push_back(S[j]); nmap.push_back(j); }
        for (int j = k_off[i]; j < k_off[i + 1]; ++j) kmap.push_back(j);
        no.push_back((int)nmap.size());
        ko.push_back((int)kmap.size());
    }
    if (!(HIP_OK(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking)) && HIP_OK(hipEventCreate(&p->ev0)) &&
          HIP_OK(hipEventCreate(&p->ev1)))) {
        batch_free(p);
        return nullptr;
    }
    if (no.size() > 1) {
        IcmHipPlanParams ip{};
        if (params) ip = *params;
        if (cfg.cap) ip.mem_budget_bytes = cfg.cap / (nl + 1);
        p->inner = batch_create((int)no.size() - 1, no.data(), Sc.data(), ko.data(), opts, &ip);
        if (!p->inner) { batch_free(p); return nullptr; }
        if (!balloc(p, &p->d_nmap, nmap.size()) || !balloc(p, &p->d_kmap, kmap.size()) ||
            !HIP_OK(hipMemcpy(p->d_nmap, nmap.data(), sizeof(int) * nmap.size(), hipMemcpyHostToDevice)) ||
            !HIP_OK(hipMemcpy(p->d_kmap, kmap.data(), sizeof(int) * kmap.size(), hipMemcpyHostToDevice))) {
            batch_free(p);
            return nullptr;
        }
        p->B = p->inner->B;
        p->qb = p->inner->qb;
    }
    if (!balloc(p, &p->d_payout, p->Ktot) || !balloc(p, &p->d_equity, p->Ntot) || !balloc(p, &p->d_vin, p->Ntot) ||
        !balloc(p, &p->d_ds, p->Ntot)) {
        batch_free(p);
        return nullptr;
    }
    if (p->verbose)
        fprintf(stderr, "[icm_hip batch] composite: %d instances in the inner batch, %d through single plans\n",
                n_inst - nl, nl);
    return p;
}

IcmHipBatchPlan *batch_create(int n_inst, const int *n_off, const double *S, const int *k_off,
                              const IcmHipOptions *opts, const IcmHipPlanParams *params) {
    if (n_inst < 1 || !n_off || !S || !k_off) { set_err("invalid arguments: need n_inst >= 1 and offsets"); return nullptr; }
    if (n_off[0] != 0 || k_off[0] != 0) { set_err("offset arrays must start at 0"); return nullptr; }
    for (int i = 0; i < n_inst; ++i) {
        const long long n = (long long)n_off[i + 1] - n_off[i], k = (long long)k_off[i + 1] - k_off[i];
        if (n < 1 || k < 1 || k > n) {
            char b[128];
            snprintf(b, sizeof b, "instance %d: need 1 <= k <= n (n=%lld, k=%lld)", i, n, k);
            set_err(b);
            return nullptr;
        }
    }
    const int Ntot = n_off[n_inst];
    for (int j = 0; j < Ntot; ++j)
        if (!(S[j] > 0.0) || !std::isfinite(S[j])) { set_err("stacks must be positive and finite"); return nullptr; }
    PlanCfg cfg;
    if (!plan_config(opts, params, cfg)) return nullptr;
    if (!ensure_device()) return nullptr;
    {
        const int Bb = cfg.B ? cfg.B : kBatchTune.B;
        std::vector<int> large;
        for (int i = 0; i < n_inst; ++i) {
            Geometry g;
            if (!make_geometry(n_off[i + 1] - n_off[i], k_off[i + 1] - k_off[i], Bb, g)) {
                set_err("invalid geometry");
                return nullptr;
            }
            for (int l = 1; l < g.L; ++l)
                if (g.F[l] > icm::kMaxLdsF) { large.push_back(i); break; }
        }
        if (!large.empty()) return composite_create(n_inst, n_off, S, k_off, opts, params, cfg, large);
    }

    IcmHipBatchPlan *p = new IcmHipBatchPlan();
    p->ninst = n_inst;
    p->Ntot = Ntot;
    p->Ktot = k_off[n_inst];
    p->n_off.assign(n_off, n_off + n_inst + 1);
    p->k_off.assign(k_off, k_off + n_inst + 1);
    p->verbose = cfg.verbose;
    p->use_graph = cfg.graphs;
    p->B = cfg.B ? cfg.B : kBatchTune.B;
    const int B = p->B;
    p->persist_grid = icm::batch_persist_grid(B);
    p->persist_ok = icm::persist_extract_blocks(B) > 0 && p->persist_grid > 0;
    p->persistent = p->persist_ok && cfg.batch_persistent;
    p->geo.resize(n_inst);
    p->boff.assign(n_inst, 0);
    p->Smax.assign(n_inst, 0.0);
    p->perm.resize(Ntot);
    p->S_sorted.resize(Ntot);
    int NB = 0;
    for (int i = 0; i < n_inst; ++i) {
        const int n = n_off[i + 1] - n_off[i], k = k_off[i + 1] - k_off[i];
        Geometry &g = p->geo[i];
        if (!make_geometry(n, k, B, g)) { set_err("invalid geometry"); delete p; return nullptr; }
        p->Lmax = std::max(p->Lmax, g.L);
        p->boff[i] = NB;
        NB += g.nb;
        // stack-descending order, index tie-break (src/gpu/gpu_api.cu:492-502)
        const int o = n_off[i];
        std::vector<int> idx(n);
        for (int j = 0; j < n; ++j) idx[j] = j;
        std::sort(idx.begin(), idx.end(), [S, o](int a, int b) {
            return S[o + a] > S[o + b] || (S[o + a] == S[o + b] && a < b);
        });
        for (int j = 0; j < n; ++j) {
            p->perm[o + j] = o + idx[j];
            p->S_sorted[o + j] = S[o + idx[j]];
        }
        p->Smax[i] = p->S_sorted[o];
    }
    p->NB = NB;
    p->blk_inst.resize(NB);
    for (int i = 0; i < n_inst; ++i)
        for (int b = 0; b < p->geo[i].nb; ++b) p->blk_inst[p->boff[i] + b] = i;

    // per-plane sizes and arena offsets (per instance, in planes of its own stride)
    const int Lm = p->Lmax;
    std::vector<long long> cplane(Lm, 0), gplane(Lm, 0);
    for (int i = 0; i < n_inst; ++i) {
        const Geometry &g = p->geo[i];
        for (int l = 1; l <= g.L - 2; ++l) cplane[l] += (long long)g.n_real[l] * g.c[l];
        for (int l = 0; l <= g.L - 2; ++l) gplane[l] += (long long)g.n_real[l] * g.G[l];
    }
    long long per_plane = (long long)NB * (B + 1);
    for (int l = 1; l < Lm; ++l) per_plane += cplane[l];
    // g: per parity, each instance's largest level of that parity (one
    // region per instance, so instances at different levels never overlap)
    long long gmax[2] = {0, 0};
    for (int i = 0; i < n_inst; ++i) {
        const Geometry &g = p->geo[i];
        long long m[2] = {0, 0};
        for (int l = 0; l <= g.L - 2; ++l) m[l & 1] = std::max(m[l & 1], (long long)g.n_real[l] * g.G[l]);
        gmax[0] += m[0];
        gmax[1] += m[1];
    }
    per_plane += gmax[0] + gmax[1];
    size_t freeb = 0, totb = 0;
    if (!HIP_OK(hipMemGetInfo(&freeb, &totb))) { delete p; return nullptr; }
    if (cfg.cap) freeb = std::min(freeb, cfg.cap);
    p->mem.cap = cfg.cap;
    const long long budget = (long long)(cfg.frac * (double)freeb);
    int qb = (int)std::max<long long>(1, budget / std::max<long long>(8 * per_plane, 1));
    qb = std::min(qb, 256);
    if (cfg.qb > 0) qb = cfg.qb;
    p->sb_max = kBatchTune.sb_max;   // one crossover per batch (tuning table)
    // arena layout for qb q-planes
    auto layout = [&](int q) {
        p->qb = q;
        p->coff.assign(Lm, std::vector<long long>(n_inst, 0));
        p->goff.assign(Lm, std::vector<long long>(n_inst, 0));
        p->coef_elems.assign(Lm, 0);
        p->g_elems[0] = p->g_elems[1] = 0;
        for (int l = 0; l < Lm; ++l) {
            long long co = 0;
            for (int i = 0; i < n_inst; ++i) {
                const Geometry &g = p->geo[i];
                if (l >= 1 && l <= g.L - 2) { p->coff[l][i] = co; co += (long long)q * g.n_real[l] * g.c[l]; }
            }
            p->coef_elems[l] = co;
        }
        // g levels of one parity share the instance's region of that parity
        long long go[2] = {0, 0};
        for (int i = 0; i < n_inst; ++i) {
            const Geometry &g = p->geo[i];
            long long m[2] = {0, 0};
            for (int l = 0; l <= g.L - 2; ++l) {
                p->goff[l][i] = go[l & 1];
                m[l & 1] = std::max(m[l & 1], (long long)q * g.n_real[l] * g.G[l]);
            }
            go[0] += m[0];
            go[1] += m[1];
        }
        p->g_elems[0] = go[0];
        p->g_elems[1] = go[1];
        const int bpw = 256 / B;
        const int wg_blocks = (NB + bpw - 1) / bpw;
        p->nchunk = std::min(std::max(1, 4096 / std::max(1, wg_blocks)), q);
    };
    auto alloc_all = [&]() {
        bool ok = balloc(p, &p->d_tw, 4097) && balloc(p, &p->d_S, Ntot) && balloc(p, &p->d_perm, Ntot) &&
                  balloc(p, &p->d_payout, p->Ktot) && balloc(p, &p->d_inst, n_inst) &&
                  balloc(p, &p->d_blk_inst, NB) && balloc(p, &p->d_blocks, (size_t)p->qb * NB * (B + 1)) &&
                  balloc(p, &p->d_g[0], p->g_elems[0]) && balloc(p, &p->d_g[1], p->g_elems[1]) &&
                  balloc(p, &p->d_partial, (size_t)p->nchunk * Ntot) && balloc(p, &p->d_acc, Ntot) &&
                  balloc(p, &p->d_equity, Ntot);
        p->d_coef.assign(Lm, nullptr);
        for (int l = 1; ok && l <= Lm - 2; ++l) ok = balloc(p, &p->d_coef[l], p->coef_elems[l]);
        if (ok && p->persist_ok)
            ok = balloc(p, &p->d_ctr, n_inst) && balloc(p, &p->d_sched, icm::kPSchedSize) &&
                 HIP_OK(hipMemset(p->d_ctr, 0, sizeof(int) * n_inst)) &&
                 HIP_OK(hipMemset(p->d_sched, 0, sizeof(int) * icm::kPSchedSize));
        return ok;
    };
    if (!(HIP_OK(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking)) && HIP_OK(hipEventCreate(&p->ev0)) &&
          HIP_OK(hipEventCreate(&p->ev1)))) {
        batch_free(p);
        return nullptr;
    }
    // out-of-memory policy of single plans: halve the q-planes per pass, at
    // most 4 times (src/gpu/gpu_memory.cu:267-279)
    for (p->retries = 0;; ++p->retries) {
        layout(qb);
        if (alloc_all()) break;
        free_buffers(p);
        if (qb > 1 && p->retries < 4) {
            qb = std::max(1, qb / 2);
            continue;
        }
        const std::string why = g_err;
        set_err("batch plan allocation failed: " + why);
        batch_free(p);
        return nullptr;
    }
    // twiddles of the in-tile transforms (every level fits one tile)
    std::vector<double2> tw(4097);
    const long double PI = 3.141592653589793238462643383279502884L;
    for (int m = 0; m < 4096; ++m) {
        const long double ang = -2.0L * PI * (long double)m / 4096.0L;
        tw[m] = make_double2((double)cosl(ang), (double)sinl(ang));
    }
    tw[4096] = make_double2(1.0, 0.0);
    p->twShift = 12;
    const bool ok = HIP_OK(hipMemcpy(p->d_tw, tw.data(), sizeof(double2) * tw.size(), hipMemcpyHostToDevice)) &&
                    HIP_OK(hipMemcpy(p->d_S, p->S_sorted.data(), sizeof(double) * Ntot, hipMemcpyHostToDevice)) &&
                    HIP_OK(hipMemcpy(p->d_perm, p->perm.data(), sizeof(int) * Ntot, hipMemcpyHostToDevice)) &&
                    HIP_OK(hipMemcpy(p->d_blk_inst, p->blk_inst.data(), sizeof(int) * NB, hipMemcpyHostToDevice));
    if (!ok) { batch_free(p); return nullptr; }
    if (p->verbose)
        fprintf(stderr, "[icm_hip batch] instances=%d players=%d blocks=%d Lmax=%d B=%d qb=%d nchunk=%d bytes=%.3f GB\n",
                n_inst, Ntot, NB, Lm, B, p->qb, p->nchunk, p->mem.bytes / 1e9);
    return p;
}

// Instance table (needs d_nodes for the current Q; B views once the adjoint
// buffers exist).
bool upload_inst(IcmHipBatchPlan *p, int Q) {
    std::vector<icm::BatchInst> inst(p->ninst);
    for (int i = 0; i < p->ninst; ++i) {
        const Geometry &g = p->geo[i];
        icm::BatchInst &I = inst[i];
        I.logv = p->d_nodes + (size_t)i * 2 * Q;
        I.w = I.logv + Q;
        I.g0 = (g.L == 1) ? icm::CoefView{p->d_payout + p->k_off[i], 0, g.k, g.k} : g_view(p, i, 0);
        I.n = g.n; I.k = g.k; I.nb = g.nb; I.L = g.L;
        I.poff = p->n_off[i];
        I.boff = p->boff[i];
        I.G0 = g.G[0];
        I.b0 = p->vjp_ready ? bv_view(p, i, 0) : icm::CoefView{};
        I.c0 = g.c[0];
    }
    return HIP_OK(hipMemcpy(p->d_inst, inst.data(), sizeof(icm::BatchInst) * inst.size(), hipMemcpyHostToDevice));
}

bool make_passes(IcmHipBatchPlan *p, int Q, bool vjp) {
    std::vector<IcmHipBatchPlan::Pass> &v = vjp ? p->vpasses : p->passes;
    free_passes(p, v);
    const int nfull = Q / p->qb, rem = Q % p->qb;
    v.resize((nfull > 0 ? 1 : 0) + (rem > 0 ? 1 : 0));
    int idx = 0;
    auto mk = [&](int nq, IcmHipBatchPlan::Pass &ps) { return vjp ? make_vjp_pass(p, nq, ps) : make_pass(p, nq, ps); };
    if (nfull > 0 && !mk(p->qb, v[idx++])) return false;
    if (rem > 0 && !mk(rem, v[idx++])) return false;
    return true;
}

// Nodes for Q (per instance: its own S_max), instance table and job tables.
bool prepare_Q(IcmHipBatchPlan *p, int Q) {
    if (p->nodes_Q == Q) return true;
    drop_graphs(p);
    free_passes(p, p->passes);
    free_passes(p, p->vpasses);
    if (p->d_nodes) {
        (void)hipFree(p->d_nodes);
        p->d_nodes = nullptr;
        p->mem.bytes -= std::min(p->mem.bytes, sizeof(double) * (size_t)p->ninst * 2 * p->nodes_alloc_Q);
    }
    p->nodes_Q = 0;
    if (!balloc(p, &p->d_nodes, (size_t)p->ninst * 2 * Q)) return false;
    p->nodes_alloc_Q = Q;
    std::vector<double> nodes((size_t)p->ninst * 2 * Q);
    for (int i = 0; i < p->ninst; ++i) {
        double *lv = nodes.data() + (size_t)i * 2 * Q;
        make_nodes(Q, p->Smax[i], lv, lv + Q);
    }
    if (!HIP_OK(hipMemcpy(p->d_nodes, nodes.data(), sizeof(double) * nodes.size(), hipMemcpyHostToDevice)) ||
        !upload_inst(p, Q) || !make_passes(p, Q, false))
        return false;
    if (p->vjp_ready && !make_passes(p, Q, true)) return false;
    p->nodes_Q = Q;
    return true;
}

// Adjoint buffers: tangent products like the forward's, B adjoint ping-pong.
bool ensure_vjp(IcmHipBatchPlan *p, int Q) {
    if (p->vjp_ready || p->composite) return true;   // a composite holds its cotangent / dS arrays from creation
    const int Lm = p->Lmax, qb = p->qb, B = p->B;
    p->bvoff.assign(Lm, std::vector<long long>(p->ninst, 0));
    for (int l = 0; l < Lm; ++l) {
        long long o = 0;
        for (int i = 0; i < p->ninst; ++i) {
            const Geometry &g = p->geo[i];
            if (l <= g.L - 2) { p->bvoff[l][i] = o; o += (long long)qb * g.n_real[l] * g.c[l]; }
        }
        p->bv_elems[l & 1] = std::max(p->bv_elems[l & 1], o);
    }
    const size_t before = p->mem.bytes;
    bool ok = balloc(p, &p->d_dblocks, (size_t)qb * p->NB * (B + 1)) && balloc(p, &p->d_bv[0], p->bv_elems[0]) &&
              balloc(p, &p->d_bv[1], p->bv_elems[1]) && balloc(p, &p->d_vin, p->Ntot) &&
              balloc(p, &p->d_v, p->Ntot) && balloc(p, &p->d_partial_ds, (size_t)p->nchunk * p->Ntot) &&
              balloc(p, &p->d_acc_ds, p->Ntot) && balloc(p, &p->d_ds, p->Ntot);
    p->d_dcoef.assign(Lm, nullptr);
    for (int l = 1; ok && l <= Lm - 2; ++l) ok = balloc(p, &p->d_dcoef[l], p->coef_elems[l]);
    if (!ok) {   // leave no partial adjoint buffers behind
        auto f = [](auto *&x) { if (x) (void)hipFree((void *)x); x = nullptr; };
        f(p->d_dblocks); for (double *&c : p->d_dcoef) f(c);
        f(p->d_bv[0]); f(p->d_bv[1]); f(p->d_vin); f(p->d_v); f(p->d_partial_ds); f(p->d_acc_ds); f(p->d_ds);
        p->mem.bytes = before;
        const std::string why = g_err;
        set_err("batch adjoint buffers: " + why);
        return false;
    }
    p->vjp_ready = true;
    if (p->nodes_Q == Q) {   // tables of this Q exist: add the B views and the adjoint jobs
        drop_graphs(p);
        return upload_inst(p, Q) && make_passes(p, Q, true);
    }
    return true;
}

// The batch as one persistent launch per q-pass (the last pass scatters).
bool enqueue_persistent(IcmHipBatchPlan *p, int Q) {
    hipStream_t s = p->stream;
    for (int q0 = 0; q0 < Q; q0 += p->qb) {
        const int nq = std::min(p->qb, Q - q0);
        const IcmHipBatchPlan::Pass &ps = (nq == p->qb) ? p->passes[0] : p->passes.back();
        icm::PersistArgs A{};
        A.items = ps.d_pitems;
        A.nitems = ps.n_pitems;
        A.qoff = ps.d_qtab;
        A.qend = ps.d_qtab + icm::kPMaxQueues + 1;
        A.nqueues = icm::kPMaxQueues;
        A.jobs = ps.d_jobs;
        A.inst = p->d_inst;
        A.blk_inst = p->d_blk_inst;
        A.S = p->d_S;
        A.blocks = p->d_blocks;
        A.NB = p->NB;
        A.nq = nq;
        A.q0 = q0;
        A.ninst = p->ninst;
        A.tw = p->d_tw;
        A.twShift = p->twShift;
        A.nchunk = ps.nch;
        A.q_per_chunk = ps.qpc;
        A.partial = p->d_partial;
        A.Ntot = p->Ntot;
        A.acc = p->d_acc;
        A.accumulate = q0 > 0 ? 1 : 0;
        A.perm = p->d_perm;
        A.equity = (q0 + nq >= Q) ? p->d_equity : nullptr;
        A.need_ex = ps.d_need;
        A.need_fin = ps.d_need + p->ninst;
        A.ctr = p->d_ctr;
        A.sched = p->d_sched;
        A.spin_limit = p->spin_limit;
        double bytes = 0.0, flops = 0.0;
        for (int l = 0; l < p->Lmax; ++l)
            bytes += ps.b_bytes[l] + ps.p_bytes[l] + ps.bs_bytes[l] + ps.ps_bytes[l],
            flops += ps.b_flops[l] + ps.p_flops[l] + ps.bs_flops[l] + ps.ps_flops[l];
        const int grid = std::min(p->persist_grid, std::max(1, ps.n_pitems));
        if (!hip_ok(icm::launch_batch_persist(p->B, A, grid, s, &p->prof, bytes, flops), "persistent batch"))
            return false;
    }
    return true;
}

// Enqueues the whole batch for Q nodes on the plan's stream (payout already
// in d_payout); leaves equities in d_equity.
bool enqueue(IcmHipBatchPlan *p, int Q) {
    if (p->persistent) return enqueue_persistent(p, Q);
    hipStream_t s = p->stream;
    const int B = p->B, NB = p->NB;
    for (int q0 = 0; q0 < Q; q0 += p->qb) {
        const int nq = std::min(p->qb, Q - q0);
        const IcmHipBatchPlan::Pass &ps = (nq == p->qb) ? p->passes[0] : p->passes.back();
        const double lb_bytes = 8.0 * ((double)p->Ntot + (double)nq * NB * (B + 1));
        if (!hip_ok(icm::launch_leaf_build_batch(B, p->d_inst, p->d_blk_inst, p->d_S, p->d_blocks, NB, nq, q0, s,
                                                 &p->prof, lb_bytes),
                    "batch leaf build"))
            return false;
        for (int l = 1; l <= p->Lmax - 2; ++l)
            if (!hip_ok(icm::launch_level_batch(true, ps.d_jobs, ps.d_items + ps.b_off[l], ps.b_cnt[l], p->d_tw,
                                                p->twShift, s, &p->prof, ps.b_bytes[l], ps.b_flops[l]),
                        "batch build level") ||
                !hip_ok(icm::launch_level_sb_batch(true, ps.d_jobs, ps.d_items + ps.bs_off[l], ps.bs_cnt[l], s,
                                                   &p->prof, ps.bs_bytes[l], ps.bs_flops[l]),
                        "batch schoolbook build level"))
                return false;
        for (int l = p->Lmax - 1; l >= 1; --l)
            if (!hip_ok(icm::launch_level_batch(false, ps.d_jobs, ps.d_items + ps.p_off[l], ps.p_cnt[l], p->d_tw,
                                                p->twShift, s, &p->prof, ps.p_bytes[l], ps.p_flops[l]),
                        "batch propagate level") ||
                !hip_ok(icm::launch_level_sb_batch(false, ps.d_jobs, ps.d_items + ps.ps_off[l], ps.ps_cnt[l], s,
                                                   &p->prof, ps.ps_bytes[l], ps.ps_flops[l]),
                        "batch schoolbook propagate level"))
                return false;
        int nch = 1, qpc = 1;
        extract_chunks(p, nq, nch, qpc);
        double g0 = 0.0;
        for (int i = 0; i < p->ninst; ++i) g0 += (double)p->geo[i].nb * p->geo[i].G[0];
        const double le_bytes = 8.0 * ((double)nq * (NB * (B + 1) + g0) + (double)p->Ntot * (1 + nch));
        if (!hip_ok(icm::launch_leaf_extract_batch(B, p->d_inst, p->d_blk_inst, p->d_S, p->d_blocks, NB, nq, q0,
                                                   nch, qpc, p->d_partial, p->Ntot, s, &p->prof, le_bytes),
                    "batch leaf extract") ||
            !hip_ok(icm::launch_finalize(p->d_partial, nch, p->Ntot, nullptr, p->d_acc, q0 > 0 ? 1 : 0, s, &p->prof),
                    "finalize"))
            return false;
    }
    return hip_ok(icm::launch_scatter(p->d_acc, p->Ntot, p->d_perm, p->d_equity, s, &p->prof), "scatter");
}

// Adjoint: gathers the cotangent (d_vin, original order), then per q-pass
// (P, dP) leaves, fused build levels, fused propagate levels (A and B),
// leaf adjoint; leaves equities in d_equity and dS in d_ds.
bool enqueue_vjp(IcmHipBatchPlan *p, int Q) {
    hipStream_t s = p->stream;
    const int B = p->B, NB = p->NB;
    if (!hip_ok(icm::launch_gather(p->d_vin, p->d_perm, p->Ntot, p->d_v, s), "gather")) return false;
    for (int q0 = 0; q0 < Q; q0 += p->qb) {
        const int nq = std::min(p->qb, Q - q0);
        const IcmHipBatchPlan::Pass &ps = (nq == p->qb) ? p->vpasses[0] : p->vpasses.back();
        const double lb_bytes = 8.0 * (2.0 * p->Ntot + 2.0 * nq * NB * (B + 1));
        if (!hip_ok(icm::launch_leaf_build_vjp_batch(B, p->d_inst, p->d_blk_inst, p->d_S, p->d_v, p->d_blocks,
                                                     p->d_dblocks, NB, nq, q0, s, &p->prof, lb_bytes),
                    "batch adjoint leaf build"))
            return false;
        for (int l = 1; l <= p->Lmax - 2; ++l)
            if (!hip_ok(icm::launch_level_t_batch(true, ps.d_jobs, ps.d_items + ps.b_off[l], ps.b_cnt[l], p->d_tw,
                                                  p->twShift, s, &p->prof, ps.b_bytes[l], ps.b_flops[l]),
                        "batch adjoint build level"))
                return false;
        for (int l = p->Lmax - 1; l >= 1; --l)
            if (!hip_ok(icm::launch_level_t_batch(false, ps.d_jobs, ps.d_items + ps.p_off[l], ps.p_cnt[l], p->d_tw,
                                                  p->twShift, s, &p->prof, ps.p_bytes[l], ps.p_flops[l]),
                        "batch adjoint propagate level"))
                return false;
        const int nchunk = std::min(p->nchunk, nq);
        const int qpc = (nq + nchunk - 1) / nchunk;
        const int nch = (nq + qpc - 1) / qpc;
        double gb = 0.0;
        for (int i = 0; i < p->ninst; ++i) gb += (double)p->geo[i].nb * (p->geo[i].G[0] + p->geo[i].c[0]);
        const double le_bytes = 8.0 * ((double)nq * (2.0 * NB * (B + 1) + gb) + (double)p->Ntot * (2 + 2 * nch));
        if (!hip_ok(icm::launch_leaf_vjp_batch(B, p->d_inst, p->d_blk_inst, p->d_S, p->d_v, p->d_blocks,
                                               p->d_dblocks, NB, nq, q0, nch, qpc, p->d_partial, p->d_partial_ds,
                                               p->Ntot, s, &p->prof, le_bytes),
                    "batch adjoint leaves") ||
            !hip_ok(icm::launch_finalize(p->d_partial, nch, p->Ntot, nullptr, p->d_acc, q0 > 0 ? 1 : 0, s, &p->prof),
                    "finalize") ||
            !hip_ok(icm::launch_finalize(p->d_partial_ds, nch, p->Ntot, nullptr, p->d_acc_ds, q0 > 0 ? 1 : 0, s,
                                         &p->prof),
                    "finalize"))
            return false;
    }
    return hip_ok(icm::launch_scatter(p->d_acc, p->Ntot, p->d_perm, p->d_equity, s, &p->prof), "scatter") &&
           hip_ok(icm::launch_scatter(p->d_acc_ds, p->Ntot, p->d_perm, p->d_ds, s, &p->prof), "scatter");
}

// Runs the batch (mode 0: forward, 1: adjoint) for Q nodes with the payout
// (and cotangent) already on the device.
int batch_run(IcmHipBatchPlan *p, int Q, IcmHipRunStats *stats, int mode = 0) {
    if (Q < 1) { set_err("Q must be positive"); return -1; }
    if (!HIP_OK(hipSetDevice(g_device))) return -1;
    if (!prepare_Q(p, Q)) return -1;
    if (mode == 1 && !ensure_vjp(p, Q)) return -1;
    hipStream_t s = p->stream;
    auto run = [&]() { return mode == 1 ? enqueue_vjp(p, Q) : enqueue(p, Q); };
    const bool graph = p->use_graph && !p->prof.enabled;
    if (graph && p->graph_Q[mode] != Q) {
        drop_graph(p, mode);
        if (!HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal))) return -1;
        const bool ok = run();
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &gr);
        if (!ok) { if (gr) (void)hipGraphDestroy(gr); return -1; }
        if (!hip_ok(e, "hipStreamEndCapture")) return -1;
        p->graph[mode] = gr;
        if (!HIP_OK(hipGraphInstantiate(&p->gexec[mode], p->graph[mode], nullptr, nullptr, 0))) {
            drop_graph(p, mode);
            return -1;
        }
        p->graph_Q[mode] = Q;
    }
    if (!HIP_OK(hipEventRecord(p->ev0, s))) return -1;
    if (graph) {
        if (!HIP_OK(hipGraphLaunch(p->gexec[mode], s))) return -1;
    } else if (!run()) {
        return -1;
    }
    if (!HIP_OK(hipEventRecord(p->ev1, s))) return -1;
    if (!HIP_OK(hipStreamSynchronize(s))) return -1;
    if (mode == 0 && p->persistent) {   // error flag of the persistent launch
        int flag = 0;
        if (!HIP_OK(hipMemcpy(&flag, p->d_sched + icm::kPSchedErr, sizeof(int), hipMemcpyDeviceToHost))) return -1;
        if (flag) {
            const int zero = 0;
            (void)hipMemcpy(p->d_sched + icm::kPSchedErr, &zero, sizeof(int), hipMemcpyHostToDevice);
            if (flag == 2) {   // an XCD queue no workgroup served: run this plan with per-level launches
                if (p->verbose) fprintf(stderr, "[icm_hip batch] persistent launch left a queue unserved; "
                                                "switching to per-level launches\n");
                p->persistent = false;
                drop_graph(p, 0);
                return batch_run(p, Q, stats, mode);
            }
            set_err("persistent batch: a work item's dependency wait passed its bound (results discarded)");
            return -1;
        }
    }
    if (p->prof.enabled) p->prof.collect();
    if (stats) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, p->ev0, p->ev1);
        memset(stats, 0, sizeof(*stats));
        stats->total_ns = (double)ms * 1e6;
        stats->peak_vram_bytes = p->mem.bytes;
        stats->engine = 1;
        stats->B = p->B;
        stats->uncached_fused_levels = 2 * p->Lmax - 3;
    }
    return 0;
}

// Composite run (payout, and cotangent for mode 1, already in d_payout /
// d_vin): the inner batch on its compacted arrays, then every large instance
// on its segment.
int composite_run(IcmHipBatchPlan *p, int Q, IcmHipRunStats *stats, int mode) {
    hipStream_t s = p->stream;
    double total_ns = 0.0;
    size_t peak = p->mem.bytes;
    IcmHipRunStats st{};
    if (p->inner) {
        IcmHipBatchPlan *b = p->inner;
        if (mode == 1 && !ensure_vjp(b, Q)) return -1;
        if (!hip_ok(icm::launch_gather(p->d_payout, p->d_kmap, b->Ktot, b->d_payout, s), "gather") ||
            (mode == 1 && !hip_ok(icm::launch_gather(p->d_vin, p->d_nmap, b->Ntot, b->d_vin, s), "gather")) ||
            !HIP_OK(hipStreamSynchronize(s)))
            return -1;
        if (batch_run(b, Q, &st, mode) != 0) return -1;
        total_ns += st.total_ns;
        peak += b->mem.bytes;
        if (!hip_ok(icm::launch_scatter(b->d_equity, b->Ntot, p->d_nmap, p->d_equity, s, nullptr), "scatter") ||
            (mode == 1 && !hip_ok(icm::launch_scatter(b->d_ds, b->Ntot, p->d_nmap, p->d_ds, s, nullptr), "scatter")))
            return -1;
    }
    if (!HIP_OK(hipStreamSynchronize(s))) return -1;
    for (size_t j = 0; j < p->large.size(); ++j) {
        const int i = p->large_ids[j];
        const double *pay = p->d_payout + p->k_off[i];
        double *eq = p->d_equity + p->n_off[i];
        const int r = mode == 1 ? icm_hip_vjp_with_plan_device(p->large[j], Q, pay, p->d_vin + p->n_off[i], eq,
                                                               p->d_ds + p->n_off[i], &st)
                                : icm_hip_equity_with_plan_device(p->large[j], Q, pay, eq, &st);
        if (r != 0) return -1;
        total_ns += st.total_ns;
        peak += (size_t)st.peak_vram_bytes;
    }
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->total_ns = total_ns;
        stats->peak_vram_bytes = peak;
        stats->engine = 1;
        stats->B = p->B;
    }
    return 0;
}

int run_any(IcmHipBatchPlan *p, int Q, IcmHipRunStats *stats, int mode = 0) {
    return p->composite ? composite_run(p, Q, stats, mode) : batch_run(p, Q, stats, mode);
}

}  // namespace

extern "C" {

IcmHipBatchPlan *icm_hip_batch_plan_create(int n_inst, const int *n_off, const double *S, const int *k_off,
                                           const IcmHipOptions *opts) {
    return batch_create(n_inst, n_off, S, k_off, opts);
}

IcmHipBatchPlan *icm_hip_batch_plan_create_ex(int n_inst, const int *n_off, const double *S, const int *k_off,
                                              const IcmHipOptions *opts, const IcmHipPlanParams *params) {
    return batch_create(n_inst, n_off, S, k_off, opts, params);
}

void icm_hip_batch_plan_destroy(IcmHipBatchPlan *plan) { batch_free(plan); }

int icm_hip_batch_equity(IcmHipBatchPlan *plan, int Q, const double *payout, double *equity,
                         IcmHipRunStats *stats) {
    if (!plan || !payout || !equity) { set_err("null argument"); return -1; }
    if (!HIP_OK(hipSetDevice(g_device))) return -1;
    if (!HIP_OK(hipMemcpyAsync(plan->d_payout, payout, sizeof(double) * plan->Ktot, hipMemcpyHostToDevice,
                               plan->stream)))
        return -1;
    if (run_any(plan, Q, stats) != 0) return -1;
    return HIP_OK(hipMemcpy(equity, plan->d_equity, sizeof(double) * plan->Ntot, hipMemcpyDeviceToHost)) ? 0 : -1;
}

int icm_hip_batch_equity_device(IcmHipBatchPlan *plan, int Q, const double *d_payout, double *d_equity,
                                IcmHipRunStats *stats) {
    if (!plan || !d_payout || !d_equity) { set_err("null argument"); return -1; }
    if (!HIP_OK(hipSetDevice(g_device))) return -1;
    if (!HIP_OK(hipMemcpyAsync(plan->d_payout, d_payout, sizeof(double) * plan->Ktot, hipMemcpyDeviceToDevice,
                               plan->stream)))
        return -1;
    if (run_any(plan, Q, stats) != 0) return -1;
    if (!HIP_OK(hipMemcpyAsync(d_equity, plan->d_equity, sizeof(double) * plan->Ntot, hipMemcpyDeviceToDevice,
                               plan->stream)))
        return -1;
    return HIP_OK(hipStreamSynchronize(plan->stream)) ? 0 : -1;
}

int icm_hip_batch_vjp(IcmHipBatchPlan *plan, int Q, const double *payout, const double *v, double *equity,
                      double *dS, IcmHipRunStats *stats) {
    if (!plan || !payout || !v) { set_err("null argument"); return -1; }
    if (!HIP_OK(hipSetDevice(g_device))) return -1;
    if (!ensure_vjp(plan, Q)) return -1;
    if (!HIP_OK(hipMemcpyAsync(plan->d_payout, payout, sizeof(double) * plan->Ktot, hipMemcpyHostToDevice,
                               plan->stream)) ||
        !HIP_OK(hipMemcpyAsync(plan->d_vin, v, sizeof(double) * plan->Ntot, hipMemcpyHostToDevice, plan->stream)))
        return -1;
    if (run_any(plan, Q, stats, 1) != 0) return -1;
    if (equity && !HIP_OK(hipMemcpy(equity, plan->d_equity, sizeof(double) * plan->Ntot, hipMemcpyDeviceToHost)))
        return -1;
    if (dS && !HIP_OK(hipMemcpy(dS, plan->d_ds, sizeof(double) * plan->Ntot, hipMemcpyDeviceToHost))) return -1;
    return 0;
}

int icm_hip_batch_vjp_device(IcmHipBatchPlan *plan, int Q, const double *d_payout, const double *d_v,
                             double *d_equity, double *d_dS, IcmHipRunStats *stats) {
    if (!plan || !d_payout || !d_v) { set_err("null argument"); return -1; }
    if (!HIP_OK(hipSetDevice(g_device))) return -1;
    if (!ensure_vjp(plan, Q)) return -1;
    hipStream_t s = plan->stream;
    if (!HIP_OK(hipMemcpyAsync(plan->d_payout, d_payout, sizeof(double) * plan->Ktot, hipMemcpyDeviceToDevice, s)) ||
        !HIP_OK(hipMemcpyAsync(plan->d_vin, d_v, sizeof(double) * plan->Ntot, hipMemcpyDeviceToDevice, s)))
        return -1;
    if (run_any(plan, Q, stats, 1) != 0) return -1;
    if (d_equity &&
        !HIP_OK(hipMemcpyAsync(d_equity, plan->d_equity, sizeof(double) * plan->Ntot, hipMemcpyDeviceToDevice, s)))
        return -1;
    if (d_dS && !HIP_OK(hipMemcpyAsync(d_dS, plan->d_ds, sizeof(double) * plan->Ntot, hipMemcpyDeviceToDevice, s)))
        return -1;
    return HIP_OK(hipStreamSynchronize(s)) ? 0 : -1;
}

int icm_hip_equity_batch(int n_inst, const int *n_off, const double *S, const int *k_off, int Q,
                         const double *payout, double *equity) {
    if (n_inst >= 1 && n_off && equity) memset(equity, 0, sizeof(double) * (size_t)std::max(0, n_off[n_inst]));
    IcmHipBatchPlan *p = batch_create(n_inst, n_off, S, k_off, nullptr);
    if (!p) return -1;
    p->use_graph = false;   // one run: recording a graph would not pay
    if (p->inner) p->inner->use_graph = false;
    const int r = icm_hip_batch_equity(p, Q, payout, equity, nullptr);
    batch_free(p);
    return r;
}

int icm_hip_batch_set_schoolbook_max(IcmHipBatchPlan *plan, int cmax) {
    if (!plan) { set_err("null plan"); return -1; }
    if (plan->composite) {
        for (IcmHipPlan *q : plan->large)
            if (icm_hip_plan_set_schoolbook_max(q, cmax) != 0) return -1;
        return plan->inner ? icm_hip_batch_set_schoolbook_max(plan->inner, cmax) : 0;
    }
    plan->sb_max = cmax < 0 ? 0 : cmax;
    drop_graphs(plan);                  // job tables change with the crossover
    free_passes(plan, plan->passes);
    free_passes(plan, plan->vpasses);
    plan->nodes_Q = 0;
    return 0;
}

int icm_hip_batch_set_persistent(IcmHipBatchPlan *plan, int on) {
    if (!plan) { set_err("null plan"); return -1; }
    if (plan->composite) return plan->inner ? icm_hip_batch_set_persistent(plan->inner, on) : 0;
    if (on && !plan->persist_ok) { set_err("the persistent batch kernel is built for block size 64 only"); return -1; }
    if (plan->persistent != (on != 0)) drop_graph(plan, 0);
    plan->persistent = on != 0;
    return 0;
}

int icm_hip_batch_is_persistent(const IcmHipBatchPlan *plan) {
    if (plan && plan->composite) return plan->inner ? icm_hip_batch_is_persistent(plan->inner) : 0;
    return plan && plan->persistent ? 1 : 0;
}

int icm_hip_batch_set_spin_limit(IcmHipBatchPlan *plan, int spins) {
    if (!plan) { set_err("null plan"); return -1; }
    if (plan->composite) return plan->inner ? icm_hip_batch_set_spin_limit(plan->inner, spins) : 0;
    if (spins < 1) { set_err("spin limit must be positive"); return -1; }
    plan->spin_limit = spins;
    drop_graph(plan, 0);   // the limit is a launch argument of the captured graph
    return 0;
}

int icm_hip_batch_set_profiling(IcmHipBatchPlan *plan, int on) {
    if (!plan) { set_err("null plan"); return -1; }
    if (plan->composite) return plan->inner ? icm_hip_batch_set_profiling(plan->inner, on) : 0;
    plan->prof.enabled = on != 0;
    plan->prof.clear();
    return 0;
}

int icm_hip_batch_kernel_count(const IcmHipBatchPlan *plan) {
    if (plan && plan->composite) return plan->inner ? icm_hip_batch_kernel_count(plan->inner) : 0;
    return plan ? (int)plan->prof.agg.size() : 0;
}

int icm_hip_batch_kernel_stat(const IcmHipBatchPlan *plan, int index, char *name, int name_len, int *launches,
                              double *total_ns, double *alg_bytes, double *alg_flops) {
    if (plan && plan->composite && plan->inner)
        return icm_hip_batch_kernel_stat(plan->inner, index, name, name_len, launches, total_ns, alg_bytes, alg_flops);
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

int icm_hip_batch_summary(const IcmHipBatchPlan *plan, int *n_inst, int *n_players, int *n_blocks, int *max_levels,
                          int *q_batch, size_t *device_bytes) {
    if (!plan) { set_err("null plan"); return -1; }
    if (plan->composite) {   // the inner batch's figures plus the single plans'
        int nb = 0, lm = 0, qb = 0;
        size_t bytes = plan->mem.bytes;
        if (plan->inner) {
            nb = plan->inner->NB;
            lm = plan->inner->Lmax;
            qb = plan->inner->qb;
            bytes += plan->inner->mem.bytes;
        }
        for (IcmHipPlan *q : plan->large) {
            IcmHipPlanSummary sm{};
            if (icm_hip_plan_summary(q, &sm) != 0) return -1;
            nb += (sm.n + sm.B - 1) / sm.B;
            lm = std::max(lm, sm.n_levels);
            qb = qb ? std::min(qb, sm.q_batch) : sm.q_batch;
            bytes += sm.planned_peak_vram_bytes;
        }
        if (n_inst) *n_inst = plan->ninst;
        if (n_players) *n_players = plan->Ntot;
        if (n_blocks) *n_blocks = nb;
        if (max_levels) *max_levels = lm;
        if (q_batch) *q_batch = qb;
        if (device_bytes) *device_bytes = bytes;
        return 0;
    }
    if (n_inst) *n_inst = plan->ninst;
    if (n_players) *n_players = plan->Ntot;
    if (n_blocks) *n_blocks = plan->NB;
    if (max_levels) *max_levels = plan->Lmax;
    if (q_batch) *q_batch = plan->qb;
    if (device_bytes) *device_bytes = plan->mem.bytes;
    return 0;
}

}  // extern "C"
