// icm_kernels.hip -- CDNA4 (gfx950) kernels of the ICM equity engine, FP64.
//
// Per quadrature node q the engine evaluates
//   Psi_i = sum_{r<k} payout[r] [x^r] prod_{j!=i} (a_j + (1 - a_j) x),
//   a_j = exp(S_j logv_q),
// through a product tree over blocks of B players, its transpose (the
// "propagate" pass, a middle product per level), and a per-block
// leave-one-out division.  Reference semantics: src/cpu/icm.c:2186-2303
// (block build, bidirectional divide), src/cpu/icm.c:1605-1819 (tree build
// and propagate), src/gpu/gpu_kernels.cu:1472-1545 (leaf extract and the
// quadrature-weighted accumulation).  All q-planes of a launch run together:
// grid.y (or grid.z) is the q-plane index.
#include "fft_lds.hpp"
#include "icm_kernels.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

namespace icm {

// 0 and 1 that a loader reads in place of a coefficient past a row's end,
// of a missing row or of the identity polynomial.  Such loaders choose an
// address and load unconditionally: with the loads under per-lane branches
// and the loaded value merged with a constant afterwards, the compiler
// waited for all earlier loads at every merge, one memory latency per point.
// (Not const: a const table would sit in the constant address space and
// turn the selected loads into flat loads.)
__device__ double kZeroOne[2] = {0.0, 1.0};
// Write-only sink for stores of lanes with nothing to store (never read).
__device__ double kSink[16];

// A pointer read from device memory (a BatchInst field) is generic to the
// compiler, and loads through it are flat loads, which count in both the
// vector-memory and the LDS counters; with one of them in flight the
// compiler waited for every outstanding load at each later wait.  Loads
// through the global address space keep the counters in order.
using gdouble = const __attribute__((address_space(1))) double;
__device__ __forceinline__ gdouble *as_global(const double *p) { return (gdouble *)p; }
using gdouble_rw = __attribute__((address_space(1))) double;
__device__ __forceinline__ gdouble_rw *as_global_rw(double *p) { return (gdouble_rw *)p; }

// Waits for every outstanding vector memory access (vmcnt 0; expcnt and
// lgkmcnt left at their maxima: simm16 0x0F70 in the gfx9 encoding).  Placed
// before a loop whose body issues loads under per-lane branches: without it
// the compiler could not tell whether loads issued before the loop were
// still in flight and waited for all loads, prefetches included, at their
// first use in every iteration.
__device__ __forceinline__ void vm_drain() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// ---------------------------------------------------------------------------
// Leaf build: one lane owns one block and multiplies its B linear factors in
// registers (triangular recurrence, src/cpu/icm.c:2196-2213).  Players past n
// are the identity factor (a = 1).  A 64 x (B+1) LDS transpose makes the
// stores of block_prods[q][b][0..B] contiguous.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double a_of(double S, double lv) {
    const double arg = S * lv;
    return (arg < -700.0) ? 0.0 : exp(arg);
}

// Coalesced store of the 64 lanes' block products (B + 1 coefficients each;
// lane i's block at out + i (B + 1)) through an LDS transpose in two halves,
// coefficients [0, H) and then [H, B + 1): a 64 x H tile is half the LDS of a
// whole one, which left room for one wave per SIMD only.
template <int B> constexpr int leaf_half() { return (B + 2) / 2; }
template <int B>
__device__ __forceinline__ void store_block_products(const double (&P)[B + 1], double *sh, double *out, int nblk) {
    constexpr int H = leaf_half<B>(), H2 = B + 1 - H;
    const int lane = threadIdx.x;
#pragma unroll
    for (int m = 0; m < H; ++m) sh[lane * H + m] = P[m];
    __syncthreads();
    for (int e = lane; e < nblk * H; e += 64) {
        const int blk = e / H;
        out[(long long)blk * (B + 1) + (e - blk * H)] = sh[e];
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < H2; ++m) sh[lane * H2 + m] = P[H + m];
    __syncthreads();
    for (int e = lane; e < nblk * H2; e += 64) {
        const int blk = e / H2;
        out[(long long)blk * (B + 1) + H + (e - blk * H2)] = sh[e];
    }
}

template <int B>
__global__ __launch_bounds__(64) void k_leaf_build(LeafJob J) {
    __shared__ double sh[64 * leaf_half<B>()];
    const int lane = threadIdx.x;
    const int qi = blockIdx.y;
    const int b0 = blockIdx.x * 64;
    const int b = b0 + lane;
    double P[B + 1];
    P[0] = 1.0;
#pragma unroll
    for (int m = 1; m <= B; ++m) P[m] = 0.0;
    if (b < J.nb) {
        const double lv = J.logv[qi];
        const long long arow = (long long)qi * J.n;
        // the players' stacks (or a values) are loaded CH at a time,
        // unconditionally (past n: kZeroOne), so each chunk waits once:
        // loaded under a per-player branch, every player waited for its
        // own load, one memory latency per player
        constexpr int CH = B < 16 ? B : 16;
        const double *src = J.a_in ? J.a_in + arow : J.S;
#pragma unroll
        for (int i0 = 0; i0 < B; i0 += CH) {
            double sv[CH];
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const int j = b * B + i0 + u;
                sv[u] = *as_global(j < J.n ? src + j : kZeroOne);
            }
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const int i = i0 + u;
                const int j = b * B + i;
                double a = 1.0;
                if (j < J.n) a = J.a_in ? sv[u] : a_of(sv[u], lv);
                const double bb = 1.0 - a;
#pragma unroll
                for (int m = i + 1; m >= 1; --m) P[m] = a * P[m] + bb * P[m - 1];
                P[0] *= a;
            }
        }
    }
    store_block_products<B>(P, sh, J.blocks + (long long)qi * J.nb * (B + 1) + (long long)b0 * (B + 1),
                            min(64, J.nb - b0));
}

// ---------------------------------------------------------------------------
// Leaf extract: Psi_j = sum_{m<pk} g_b[m] (P_b / (a_j + (1-a_j)x))[m] with the
// bidirectional division of src/cpu/icm.c:2244-2299 (forward if a > 0.5,
// backward if 1e-15 < a <= 0.5, shift otherwise), then the quadrature weight
// w_q S_j a_j exp(-logv_q) (non-finite -> 0, src/cpu/icm.c:2584-2589).
// One lane per player; the lane loops over its chunk of q-planes and keeps the
// sum in a register, so the q-sum order is fixed (no atomics).
// ---------------------------------------------------------------------------
// Psi of one player of a block of bsize players: sum_{m<pk} g[m] (P/f)[m],
// f = a + (1-a) x, dividing in the stable direction.
template <int B>
__device__ __forceinline__ double leaf_psi(const double *P, const double *g, double a, int bsize, int pk) {
    const double bm = 1.0 - a;
    double eq = 0.0;
    if (a > 0.5) {
        const double ia = 1.0 / a, c = -bm * ia;
        double qv = P[0] * ia;
        eq = g[0] * qv;
        for (int m = 1; m < pk; ++m) {
            qv = c * qv + P[m] * ia;
            eq += g[m] * qv;
        }
    } else if (a > 1e-15) {
        const double ib = 1.0 / bm, c = -a * ib;
        double qv = P[bsize] * ib;
        if (bsize - 1 < pk) eq += g[bsize - 1] * qv;
        // the quotient runs down from the top; only m < pk meets g (two
        // branch-free loops instead of a test per step)
        int m = bsize - 2;
        for (; m >= pk; --m) qv = c * qv + P[m + 1] * ib;
        for (; m >= 0; --m) {
            qv = c * qv + P[m + 1] * ib;
            eq += g[m] * qv;
        }
    } else {
        for (int m = 0; m < pk; ++m) eq += g[m] * P[m + 1];
    }
    return eq;
}

// leaf_psi for PPL players of one block at once: each coefficient read from
// LDS feeds PPL recurrences.  Per player the operations and their order are
// those of leaf_psi (the branch not taken by a player is computed for the
// others and discarded).
template <int B, int PPL>
__device__ __forceinline__ void leaf_psi_n(const double *P, const double *g, const double *a, int bsize, int pk,
                                           double *eq) {
    bool fw = false, bw = false, zr = false;
#pragma unroll
    for (int u = 0; u < PPL; ++u) {
        fw |= a[u] > 0.5;
        bw |= (a[u] <= 0.5) && (a[u] > 1e-15);
        zr |= a[u] <= 1e-15;
    }
    double ef[PPL], eb[PPL], ez = 0.0;
    if (fw) {
        double ia[PPL], c[PPL], qv[PPL];
#pragma unroll
        for (int u = 0; u < PPL; ++u) {
            const double bm = 1.0 - a[u];
            ia[u] = 1.0 / a[u];
            c[u] = -bm * ia[u];
            qv[u] = P[0] * ia[u];
            ef[u] = g[0] * qv[u];
        }
        for (int m = 1; m < pk; ++m) {
            const double pm = P[m], gm = g[m];
#pragma unroll
            for (int u = 0; u < PPL; ++u) {
                qv[u] = c[u] * qv[u] + pm * ia[u];
                ef[u] += gm * qv[u];
            }
        }
    }
    if (bw) {
        double ib[PPL], c[PPL], qv[PPL];
#pragma unroll
        for (int u = 0; u < PPL; ++u) {
            const double bm = 1.0 - a[u];
            ib[u] = 1.0 / bm;
            c[u] = -a[u] * ib[u];
            qv[u] = P[bsize] * ib[u];
            eb[u] = 0.0;
            if (bsize - 1 < pk) eb[u] += g[bsize - 1] * qv[u];
        }
        int m = bsize - 2;
        for (; m >= pk; --m) {
            const double pm = P[m + 1];
#pragma unroll
            for (int u = 0; u < PPL; ++u) qv[u] = c[u] * qv[u] + pm * ib[u];
        }
        for (; m >= 0; --m) {
            const double pm = P[m + 1], gm = g[m];
#pragma unroll
            for (int u = 0; u < PPL; ++u) {
                qv[u] = c[u] * qv[u] + pm * ib[u];
                eb[u] += gm * qv[u];
            }
        }
    }
    if (zr)
        for (int m = 0; m < pk; ++m) ez += g[m] * P[m + 1];
#pragma unroll
    for (int u = 0; u < PPL; ++u) eq[u] = a[u] > 0.5 ? ef[u] : (a[u] > 1e-15 ? eb[u] : ez);
}

// Leaf extract with PPL players of one block per lane (players li, li + B/PPL,
// ...): B/PPL lanes per block, 256 PPL / B blocks per workgroup.  LDS reads
// per player drop by PPL (the LDS-staged broadcast reads of the block
// products and g bound the one-player kernel).  Same results, bit for bit,
// as k_leaf_extract.
template <int B, int PPL>
__global__ __launch_bounds__(256) void k_leaf_extract_n(LeafJob J) {
    constexpr int LPB = B / PPL;                 // lanes per block
    constexpr int BPW = 256 / LPB;               // blocks per workgroup
    constexpr int NP = BPW * (B + 1);            // block coefficients per q-plane
    constexpr int NPT = (NP + 255) / 256;        // of them per thread
    constexpr int NG = BPW * B;                  // g entries per q-plane (PPL per thread)
    static_assert(LPB * PPL == B && BPW * LPB == 256, "lanes must tile the blocks");
    __shared__ double shP[2][NP];
    __shared__ double shG[2][NG];
    const int t = threadIdx.x;
    const int lb = t / LPB, li = t % LPB;
    const int sfirst = blockIdx.x * BPW;
    auto block_of = [&](int slot) { return slot < J.nbl ? (J.blist ? J.blist[slot] : slot) : -1; };
    const int b = block_of(sfirst + lb);
    const int chunk = blockIdx.y;
    const int q0 = chunk * J.q_per_chunk;
    const int q1 = min(q0 + J.q_per_chunk, J.nq);
    const int bsize = (b >= 0) ? min(B, J.n - b * B) : 0;
    const int pk = min(J.G0, bsize);
    int jj[PPL];
    bool valid[PPL], act[PPL];
    double Sj[PPL];
    bool any = false;
#pragma unroll
    for (int u = 0; u < PPL; ++u) {
        jj[u] = b * B + li + u * LPB;
        valid[u] = (b >= 0) && (jj[u] < J.n);
        Sj[u] = valid[u] ? J.S[jj[u]] : 0.0;
        act[u] = valid[u] && (!J.active || J.active[jj[u]]);
        any |= act[u];
    }
    const long long bplane = (long long)J.nb * (B + 1);
    int bP[NPT], mP[NPT], bG[PPL], mG[PPL];
#pragma unroll
    for (int i = 0; i < NPT; ++i) {
        const int e = t + i * 256;
        bP[i] = (e < NP) ? block_of(sfirst + e / (B + 1)) : -1;
        mP[i] = e % (B + 1);
    }
#pragma unroll
    for (int i = 0; i < PPL; ++i) {
        const int e = t + i * 256;
        bG[i] = block_of(sfirst + e / B);
        mG[i] = e % B;
        if (mG[i] >= J.G0) bG[i] = -1;
    }
    double pr[NPT], gr[PPL];
    auto fetch = [&](int q) {
#pragma unroll
        for (int i = 0; i < NPT; ++i)
            pr[i] = (bP[i] >= 0) ? J.blocks[q * bplane + (long long)bP[i] * (B + 1) + mP[i]] : 0.0;
#pragma unroll
        for (int i = 0; i < PPL; ++i)
            gr[i] = (bG[i] >= 0) ? J.g[q * J.g_plane + (long long)bG[i] * J.g_stride + mG[i]] : 0.0;
    };
    auto stash = [&](int buf) {
#pragma unroll
        for (int i = 0; i < NPT; ++i)
            if (t + i * 256 < NP) shP[buf][t + i * 256] = pr[i];
#pragma unroll
        for (int i = 0; i < PPL; ++i) shG[buf][t + i * 256] = gr[i];
    };
    if (q0 < q1) {
        fetch(q0);
        stash(0);
    }
    __syncthreads();
    double acc[PPL];
#pragma unroll
    for (int u = 0; u < PPL; ++u) acc[u] = 0.0;
    for (int q = q0; q < q1; ++q) {
        const int cur = (q - q0) & 1;
        const bool more = q + 1 < q1;
        if (more) fetch(q + 1);
        if (any) {
            const double lv = J.logv[q];
            double a[PPL], eq[PPL];
#pragma unroll
            for (int u = 0; u < PPL; ++u)
                a[u] = !act[u] ? 0.75 : (J.a_in ? J.a_in[(long long)q * J.n + jj[u]] : a_of(Sj[u], lv));
            leaf_psi_n<B, PPL>(shP[cur] + lb * (B + 1), shG[cur] + lb * B, a, bsize, pk, eq);
            const double el = exp(-lv), wq = J.wq[q];
#pragma unroll
            for (int u = 0; u < PPL; ++u) {
                if (!act[u]) continue;
                if (J.inner_out) J.inner_out[(long long)q * J.n + jj[u]] = eq[u];
                double pw = wq * Sj[u] * a[u] * el;
                if (!isfinite(pw)) pw = 0.0;
                acc[u] += pw * eq[u];
            }
        }
        if (more) stash(cur ^ 1);
        __syncthreads();
    }
    if (J.partial) {
#pragma unroll
        for (int u = 0; u < PPL; ++u)
            if (valid[u]) J.partial[(long long)chunk * J.n + jj[u]] = act[u] ? acc[u] : 0.0;
    }
}

// Players per lane of the B = 64 leaf extracts: 2, or 1 when
// ICM_HIP_LEAF_PPL=1 (read at each launch; a captured batch graph keeps the
// choice of its capture) to compare with the one-player kernels.  Four per
// lane measured slower (171 VGPRs, two waves per SIMD).
inline int leaf_ppl() {
    const char *e = getenv("ICM_HIP_LEAF_PPL");
    return (e && atoi(e) == 1) ? 1 : 2;
}

template <int B>
__global__ __launch_bounds__(256) void k_leaf_extract(LeafJob J) {
    constexpr int BPW = 256 / B;
    constexpr int NP = BPW * (B + 1);          // 256 + BPW block coefficients per q-plane
    static_assert(BPW * B == 256, "one g entry per thread");
    __shared__ double shP[2][NP];
    __shared__ double shG[2][256];
    const int t = threadIdx.x;
    const int lb = t / B, li = t % B;
    const int sfirst = blockIdx.x * BPW;   // first block slot of this workgroup
    auto block_of = [&](int slot) { return slot < J.nbl ? (J.blist ? J.blist[slot] : slot) : -1; };
    const int b = block_of(sfirst + lb);
    const int j = b * B + li;
    const int chunk = blockIdx.y;
    const int q0 = chunk * J.q_per_chunk;
    const int q1 = min(q0 + J.q_per_chunk, J.nq);
    const bool valid = (b >= 0) && (j < J.n);
    const int bsize = (b >= 0) ? min(B, J.n - b * B) : 0;
    const int pk = min(J.G0, bsize);
    const double Sj = valid ? J.S[j] : 0.0;
    const bool act = valid && (!J.active || J.active[j]);
    const long long bplane = (long long)J.nb * (B + 1);
    // this thread's share of a q-plane's operands: P elements t and t + 256, g element t
    const int e1 = t + 256;
    const bool has1 = e1 < NP;
    const int bP0 = block_of(sfirst + t / (B + 1)), mP0 = t % (B + 1);
    const int bP1 = has1 ? block_of(sfirst + e1 / (B + 1)) : -1, mP1 = e1 % (B + 1);
    const int bG = block_of(sfirst + t / B), mG = t % B;
    const bool gok = bG >= 0 && mG < J.G0;
    double p0 = 0.0, p1 = 0.0, g0 = 0.0;
    auto fetch = [&](int q) {
        p0 = (bP0 >= 0) ? J.blocks[q * bplane + (long long)bP0 * (B + 1) + mP0] : 0.0;
        p1 = (bP1 >= 0) ? J.blocks[q * bplane + (long long)bP1 * (B + 1) + mP1] : 0.0;
        g0 = gok ? J.g[q * J.g_plane + (long long)bG * J.g_stride + mG] : 0.0;
    };
    auto stash = [&](int buf) {
        shP[buf][t] = p0;
        if (has1) shP[buf][e1] = p1;
        shG[buf][t] = g0;
    };
    if (q0 < q1) {
        fetch(q0);
        stash(0);
    }
    __syncthreads();
    double acc = 0.0;
    // double-buffered: plane q+1's loads are in flight while plane q is computed
    for (int q = q0; q < q1; ++q) {
        const int cur = (q - q0) & 1;
        const bool more = q + 1 < q1;
        if (more) fetch(q + 1);
        if (act) {
            const double lv = J.logv[q];
            const double a = J.a_in ? J.a_in[(long long)q * J.n + j] : a_of(Sj, lv);
            const double eq = leaf_psi<B>(shP[cur] + lb * (B + 1), shG[cur] + lb * B, a, bsize, pk);
            if (J.inner_out) J.inner_out[(long long)q * J.n + j] = eq;
            double pw = J.wq[q] * Sj * a * exp(-lv);
            if (!isfinite(pw)) pw = 0.0;
            acc += pw * eq;
        }
        if (more) stash(cur ^ 1);
        __syncthreads();
    }
    if (valid && J.partial) J.partial[(long long)chunk * J.n + j] = act ? acc : 0.0;
}

// ---------------------------------------------------------------------------
// Ragged-batch leaves.  Blocks of all instances are numbered globally
// (instance i owns blocks boff .. boff + nb - 1, players poff .. poff + n - 1
// of the concatenated sorted stacks); blk_inst maps a block to its instance.
// Block products live at blocks[q][global block][B+1].
// ---------------------------------------------------------------------------
template <int B>
__global__ __launch_bounds__(64) void k_leaf_build_batch(const BatchInst *__restrict__ inst,
                                                         const int *__restrict__ blk_inst,
                                                         const double *__restrict__ S, double *blocks, int NB,
                                                         int q0) {
    __shared__ double sh[64 * leaf_half<B>()];
    const int lane = threadIdx.x;
    const int qi = blockIdx.y;
    const int b0 = blockIdx.x * 64;
    const int gb = b0 + lane;
    double P[B + 1];
    P[0] = 1.0;
#pragma unroll
    for (int m = 1; m <= B; ++m) P[m] = 0.0;
    if (gb < NB) {
        const BatchInst &I = inst[blk_inst[gb]];
        const double lv = I.logv[q0 + qi];
        const int b = gb - I.boff;
        // stacks loaded CH players at a time, unconditionally (as k_leaf_build)
        constexpr int CH = B < 16 ? B : 16;
#pragma unroll
        for (int i0 = 0; i0 < B; i0 += CH) {
            double sv[CH];
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const int j = b * B + i0 + u;
                sv[u] = *as_global(j < I.n ? S + I.poff + j : kZeroOne);
            }
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const int i = i0 + u;
                const double a = (b * B + i < I.n) ? a_of(sv[u], lv) : 1.0;
                const double bb = 1.0 - a;
#pragma unroll
                for (int m = i + 1; m >= 1; --m) P[m] = a * P[m] + bb * P[m - 1];
                P[0] *= a;
            }
        }
    }
    store_block_products<B>(P, sh, blocks + (long long)qi * NB * (B + 1) + (long long)b0 * (B + 1), min(64, NB - b0));
}

// Ragged-batch leaf extract with PPL players of one block per lane (as
// k_leaf_extract_n); the same results, bit for bit, as k_leaf_extract_batch.
template <int B, int PPL> struct LeafExN {
    static constexpr int LPB = B / PPL;          // lanes per block
    static constexpr int BPW = 256 / LPB;        // blocks per 256 threads
    static constexpr int NP = BPW * (B + 1);     // block-product doubles per q-plane
    static constexpr int NG = BPW * B;           // g doubles per q-plane
    static_assert(LPB * PPL == B && BPW * LPB == 256, "lanes must tile the blocks");
};

// Body of the batch extract over blocks gfirst .. gfirst + BPW - 1 (those
// below gend; NB is the block count of the arrays) and q-chunk `chunk`, run
// by 256 threads (t = 0 .. 255) with LDS shP[2][NP], shG[2][NG].
// Its barrier count depends on the chunk only, so two halves of a 512-thread
// workgroup may run it on neighbouring block groups of one chunk.
template <int B, int PPL>
__device__ __forceinline__ void leaf_extract_batch_n_body(int t, int gfirst, int chunk,
                                                          const BatchInst *__restrict__ inst,
                                                          const int *__restrict__ blk_inst,
                                                          const double *__restrict__ S,
                                                          const double *__restrict__ blocks, int NB, int gend,
                                                          int nq, int q0, int q_per_chunk, double *partial, int Ntot,
                                                          double (*shP)[LeafExN<B, PPL>::NP],
                                                          double (*shG)[LeafExN<B, PPL>::NG]) {
    constexpr int LPB = LeafExN<B, PPL>::LPB;
    constexpr int NP = LeafExN<B, PPL>::NP;
    constexpr int NPT = (NP + 255) / 256;
    const int lb = t / LPB, li = t % LPB;
    const int gb = gfirst + lb;
    const int qa = chunk * q_per_chunk;
    const int qe = min(qa + q_per_chunk, nq);
    const BatchInst *I = (gb < gend) ? &inst[blk_inst[gb]] : nullptr;
    const int b = I ? gb - I->boff : 0;
    const int bsize = I ? min(B, I->n - b * B) : 0;
    const int pk = I ? min(I->G0, bsize) : 0;
    int jj[PPL];
    bool valid[PPL];
    double Sj[PPL];
    bool any = false;
#pragma unroll
    for (int u = 0; u < PPL; ++u) {
        jj[u] = b * B + li + u * LPB;
        valid[u] = I && jj[u] < I->n;
        Sj[u] = valid[u] ? S[I->poff + jj[u]] : 0.0;
        any |= valid[u];
    }
    const long long bplane = (long long)NB * (B + 1);
    int gP[NPT], mP[NPT];
#pragma unroll
    for (int i = 0; i < NPT; ++i) {
        const int e = t + i * 256;
        gP[i] = (e < NP && gfirst + e / (B + 1) < gend) ? gfirst + e / (B + 1) : NB;
        mP[i] = e % (B + 1);
    }
    gdouble *gbase[PPL];
    long long gplane[PPL];
#pragma unroll
    for (int i = 0; i < PPL; ++i) {
        const int e = t + i * 256;
        const int gG = gfirst + e / B, mG = e % B;
        gbase[i] = nullptr;
        gplane[i] = 0;
        if (gG < gend) {
            const BatchInst &I2 = inst[blk_inst[gG]];
            if (mG < I2.G0) {
                gbase[i] = as_global(I2.g0.base + (long long)(gG - I2.boff) * I2.g0.stride + mG);
                gplane[i] = I2.g0.plane;
            }
        }
    }
    // The block's node and weight of a q-plane are fetched with that plane's
    // block products, ahead of them: read inside the loop, after the next
    // plane's loads, they waited for those loads (in-order vector memory
    // counter), one memory latency per q-plane.  Lanes without a block read
    // the 0 of kZeroOne.
    gdouble *nodes = as_global(I ? I->logv + q0 : kZeroOne);
    gdouble *wts = as_global(I ? I->w + q0 : kZeroOne);
    const int qmul = I ? 1 : 0;
    double pr[NPT], gr[PPL], lvn = 0.0, wn = 0.0;
    auto fetch = [&](int q) {
        lvn = nodes[q * qmul];
        wn = wts[q * qmul];
#pragma unroll
        for (int i = 0; i < NPT; ++i)
            pr[i] = (gP[i] < NB) ? blocks[q * bplane + (long long)gP[i] * (B + 1) + mP[i]] : 0.0;
#pragma unroll
        for (int i = 0; i < PPL; ++i) gr[i] = gbase[i] ? gbase[i][q * gplane[i]] : 0.0;
    };
    auto stash = [&](int buf) {
#pragma unroll
        for (int i = 0; i < NPT; ++i)
            if (t + i * 256 < NP) shP[buf][t + i * 256] = pr[i];
#pragma unroll
        for (int i = 0; i < PPL; ++i) shG[buf][t + i * 256] = gr[i];
    };
    double lv = 0.0, wq = 0.0;
    if (qa < qe) {
        fetch(qa);
        stash(0);
        lv = lvn;
        wq = wn;
    }
    vm_drain();
    __syncthreads();
    double acc[PPL];
#pragma unroll
    for (int u = 0; u < PPL; ++u) acc[u] = 0.0;
    for (int q = qa; q < qe; ++q) {
        const int cur = (q - qa) & 1;
        const bool more = q + 1 < qe;
        if (more) fetch(q + 1);
        if (any) {
            double a[PPL], eq[PPL];
#pragma unroll
            for (int u = 0; u < PPL; ++u) a[u] = valid[u] ? a_of(Sj[u], lv) : 0.75;
            leaf_psi_n<B, PPL>(shP[cur] + lb * (B + 1), shG[cur] + lb * B, a, bsize, pk, eq);
            const double el = exp(-lv);
#pragma unroll
            for (int u = 0; u < PPL; ++u) {
                double pw = wq * Sj[u] * a[u] * el;
                if (!isfinite(pw)) pw = 0.0;
                acc[u] += pw * eq[u];
            }
        }
        if (more) {
            stash(cur ^ 1);
            lv = lvn;
            wq = wn;
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < PPL; ++u)
        if (valid[u]) partial[(long long)chunk * Ntot + I->poff + jj[u]] = acc[u];
}

template <int B, int PPL>
__global__ __launch_bounds__(256, 4) void k_leaf_extract_batch_n(const BatchInst *__restrict__ inst,
                                                                 const int *__restrict__ blk_inst,
                                                                 const double *__restrict__ S,
                                                                 const double *__restrict__ blocks, int NB, int nq,
                                                                 int q0, int q_per_chunk, double *partial, int Ntot) {
    using X = LeafExN<B, PPL>;
    __shared__ double shP[2][X::NP];
    __shared__ double shG[2][X::NG];
    leaf_extract_batch_n_body<B, PPL>(threadIdx.x, blockIdx.x * X::BPW, blockIdx.y, inst, blk_inst, S, blocks, NB, NB,
                                      nq, q0, q_per_chunk, partial, Ntot, shP, shG);
}

template <int B>
__global__ __launch_bounds__(256) void k_leaf_extract_batch(const BatchInst *__restrict__ inst,
                                                            const int *__restrict__ blk_inst,
                                                            const double *__restrict__ S,
                                                            const double *__restrict__ blocks, int NB, int nq,
                                                            int q0, int q_per_chunk, double *partial, int Ntot) {
    constexpr int BPW = 256 / B;
    constexpr int NP = BPW * (B + 1);
    static_assert(BPW * B == 256, "one g entry per thread");
    __shared__ double shP[2][NP];
    __shared__ double shG[2][256];
    const int t = threadIdx.x;
    const int lb = t / B, li = t % B;
    const int gfirst = blockIdx.x * BPW;
    const int gb = gfirst + lb;
    const int chunk = blockIdx.y;
    const int qa = chunk * q_per_chunk;
    const int qe = min(qa + q_per_chunk, nq);
    const BatchInst *I = (gb < NB) ? &inst[blk_inst[gb]] : nullptr;
    const int b = I ? gb - I->boff : 0;
    const int j = b * B + li;
    const bool valid = I && j < I->n;
    const int bsize = I ? min(B, I->n - b * B) : 0;
    const int pk = I ? min(I->G0, bsize) : 0;
    const double Sj = valid ? S[I->poff + j] : 0.0;
    const long long bplane = (long long)NB * (B + 1);
    // this thread's share of a q-plane's operands: P elements t and t + 256, g element t
    const int e1 = t + 256;
    const bool has1 = e1 < NP;
    const int gP0 = gfirst + t / (B + 1), mP0 = t % (B + 1);
    const int gP1 = has1 ? gfirst + e1 / (B + 1) : NB, mP1 = e1 % (B + 1);
    const int gG = gfirst + t / B, mG = t % B;
    gdouble *gbase = nullptr;
    long long gplane = 0;
    if (gG < NB) {
        const BatchInst &I2 = inst[blk_inst[gG]];
        if (mG < I2.G0) {
            gbase = as_global(I2.g0.base + (long long)(gG - I2.boff) * I2.g0.stride + mG);
            gplane = I2.g0.plane;
        }
    }
    // node and weight fetched with the plane's operands, ahead of them (as
    // leaf_extract_batch_n_body)
    gdouble *nodes = as_global(I ? I->logv + q0 : kZeroOne);
    gdouble *wts = as_global(I ? I->w + q0 : kZeroOne);
    const int qmul = I ? 1 : 0;
    double p0 = 0.0, p1 = 0.0, g0 = 0.0, lvn = 0.0, wn = 0.0;
    auto fetch = [&](int q) {
        lvn = nodes[q * qmul];
        wn = wts[q * qmul];
        p0 = (gP0 < NB) ? blocks[q * bplane + (long long)gP0 * (B + 1) + mP0] : 0.0;
        p1 = (gP1 < NB) ? blocks[q * bplane + (long long)gP1 * (B + 1) + mP1] : 0.0;
        g0 = gbase ? gbase[q * gplane] : 0.0;
    };
    auto stash = [&](int buf) {
        shP[buf][t] = p0;
        if (has1) shP[buf][e1] = p1;
        shG[buf][t] = g0;
    };
    double lv = 0.0, wq = 0.0;
    if (qa < qe) {
        fetch(qa);
        stash(0);
        lv = lvn;
        wq = wn;
    }
    vm_drain();
    __syncthreads();
    double acc = 0.0;
    for (int q = qa; q < qe; ++q) {
        const int cur = (q - qa) & 1;
        const bool more = q + 1 < qe;
        if (more) fetch(q + 1);
        if (valid) {
            const double a = a_of(Sj, lv);
            const double eq = leaf_psi<B>(shP[cur] + lb * (B + 1), shG[cur] + lb * B, a, bsize, pk);
            double pw = wq * Sj * a * exp(-lv);
            if (!isfinite(pw)) pw = 0.0;
            acc += pw * eq;
        }
        if (more) {
            stash(cur ^ 1);
            lv = lvn;
            wq = wn;
        }
        __syncthreads();
    }
    if (valid) partial[(long long)chunk * Ntot + I->poff + j] = acc;
}

// ---------------------------------------------------------------------------
// Adjoint leaves.  With tangent weights c_j = v_j w_q S_j a_j exp(-logv_q)
// (the coefficient of Psi_j in v.E) every factor f_j = a_j + (1-a_j) x gets
// the tangent c_j (direction a_j, b_j held), so a block carries (P_b, dP_b)
// with dP <- dP f + c P, P <- P f.
// ---------------------------------------------------------------------------
template <int B>
__global__ __launch_bounds__(64) void k_leaf_build_vjp(VjpLeafJob V) {
    const LeafJob &J = V.L;
    __shared__ double sh[64 * (B + 1)];
    const int lane = threadIdx.x;
    const int qi = blockIdx.y;
    const int b0 = blockIdx.x * 64;
    const int b = b0 + lane;
    double P[B + 1], D[B + 1];
    P[0] = 1.0;
    D[0] = 0.0;
#pragma unroll
    for (int m = 1; m <= B; ++m) { P[m] = 0.0; D[m] = 0.0; }
    if (b < J.nb) {
        const double lv = J.logv[qi], wq = J.wq[qi];
        const double ivv = exp(-lv);
        const long long arow = (long long)qi * J.n;
        // operands loaded CH players at a time, unconditionally (as
        // k_leaf_build): stack, a value (with a_in) and c_in or v
        constexpr int CH = 8;
        static_assert(B % CH == 0, "whole chunks");
#pragma unroll
        for (int i0 = 0; i0 < B; i0 += CH) {
            double sv[CH], av[CH], cv[CH];
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const int j = b * B + i0 + u;
                const bool ok = j < J.n;
                sv[u] = *as_global(ok ? J.S + j : kZeroOne);
                av[u] = J.a_in ? *as_global(ok ? J.a_in + arow + j : kZeroOne) : 0.0;
                cv[u] = *as_global(ok ? (V.c_in ? V.c_in + arow + j : V.v + j) : kZeroOne);
            }
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const int i = i0 + u;
                const int j = b * B + i;
                double a = 1.0, c = 0.0;
                if (j < J.n) {
                    a = J.a_in ? av[u] : a_of(sv[u], lv);
                    if (V.c_in) {
                        c = cv[u];
                    } else {
                        double pw = wq * sv[u] * a * ivv;
                        if (!isfinite(pw)) pw = 0.0;
                        c = cv[u] * pw;
                    }
                }
                const double bb = 1.0 - a;
#pragma unroll
                for (int m = i + 1; m >= 1; --m) {
                    D[m] = a * D[m] + bb * D[m - 1] + c * P[m];
                    P[m] = a * P[m] + bb * P[m - 1];
                }
                D[0] = a * D[0] + c * P[0];
                P[0] *= a;
            }
        }
    }
    const int nblk = min(64, J.nb - b0);
    const long long off = (long long)qi * J.nb * (B + 1) + (long long)b0 * (B + 1);
#pragma unroll
    for (int m = 0; m <= B; ++m) sh[lane * (B + 1) + m] = P[m];
    __syncthreads();
    for (int e = lane; e < nblk * (B + 1); e += 64) J.blocks[off + e] = sh[e];
    __syncthreads();
#pragma unroll
    for (int m = 0; m <= B; ++m) sh[lane * (B + 1) + m] = D[m];
    __syncthreads();
    for (int e = lane; e < nblk * (B + 1); e += 64) V.dblocks[off + e] = sh[e];
}

// One player's adjoint at the leaves (A = forward g, Bv = adjoint of P_b,
// both zero past their lengths; A has B+1 entries, Bv B+2):
//   Q = P_b / f_j,  W = (dP_b - c_j Q) / f_j   (same division branch)
//   psi = sum_{m<pk} A[m] Q[m]
//   h   = sum_m Q[m] (Bv[m] - Bv[m+1]) + sum_m W[m] (A[m] - A[m+1])
template <int B>
__device__ __forceinline__ void leaf_vjp_player(const double *P, const double *D, const double *A, const double *Bv,
                                                double a, double c, int bsize, int pk, double &psi, double &h) {
    const double bm = 1.0 - a;
    psi = 0.0;
    h = 0.0;
    // Each direction runs its steps in order as a few branch-free loops: the
    // psi term only for m < pk, the W recurrence only where W exists.
    using Y = std::true_type;
    using N = std::false_type;
    if (a > 0.5) {
        const double ia = 1.0 / a, r = -bm * ia;
        double qv = 0.0, wv = 0.0;
        auto step = [&](int m, auto psi_on, auto w_on) {
            qv = P[m] * ia + r * qv;
            if constexpr (decltype(psi_on)::value) psi += A[m] * qv;
            h += qv * (Bv[m] - Bv[m + 1]);
            if constexpr (decltype(w_on)::value) {
                wv = (D[m] - c * qv) * ia + r * wv;
                h += wv * (A[m] - A[m + 1]);
            }
        };
        int m = 0;
        for (; m < pk && m <= bsize - 2; ++m) step(m, Y{}, Y{});
        for (; m < pk; ++m) step(m, Y{}, N{});            // m = bsize - 1 when pk = bsize
        for (; m <= bsize - 2; ++m) step(m, N{}, Y{});
        for (; m < bsize; ++m) step(m, N{}, N{});
    } else if (a > 1e-15) {
        const double ib = 1.0 / bm, r = -a * ib;
        double qv = 0.0, wv = 0.0;
        auto step = [&](int m, auto psi_on, auto w_on) {
            qv = P[m + 1] * ib + r * qv;
            if constexpr (decltype(psi_on)::value) psi += A[m] * qv;
            h += qv * (Bv[m] - Bv[m + 1]);
            if constexpr (decltype(w_on)::value) {
                wv = (D[m] - c * qv) * ib + r * wv;     // W[m-1]
                h += wv * (A[m - 1] - A[m]);
            }
        };
        int m = bsize - 1;   // the top step starts the quotient: Q[bsize-1] = P[bsize] / (1 - a)
        qv = P[bsize] * ib;
        if (m < pk) psi += A[m] * qv;
        h += qv * (Bv[m] - Bv[m + 1]);
        if (m >= 1) {
            wv = (D[m] - c * qv) * ib + r * wv;
            h += wv * (A[m - 1] - A[m]);
        }
        --m;
        for (; m >= pk && m >= 1; --m) step(m, N{}, Y{});
        for (; m >= pk; --m) step(m, N{}, N{});            // m = 0 when pk = 0
        for (; m >= 1; --m) step(m, Y{}, Y{});
        for (; m >= 0; --m) step(m, Y{}, N{});
    } else {
        for (int m = 0; m < bsize; ++m) {
            const double qv = P[m + 1];
            if (m < pk) psi += A[m] * qv;
            h += qv * (Bv[m] - Bv[m + 1]);
            if (m <= bsize - 2) {
                const double wv = D[m + 1] - c * (m + 2 <= B ? P[m + 2] : 0.0);
                h += wv * (A[m] - A[m + 1]);
            }
        }
    }
}

// leaf_vjp_player for PPL players of one block at once, with the
// differences dA[m] = A[m] - A[m+1] and dB[m] = Bv[m] - Bv[m+1] staged
// beside A (the same values leaf_vjp_player forms per step).  Per player the
// operations and their order are those of leaf_vjp_player; a branch not taken
// by a player is computed for the others and discarded.
template <int B, int PPL>
__device__ __forceinline__ void leaf_vjp_n(const double *P, const double *D, const double *A, const double *dA,
                                           const double *dB, const double *a, const double *c, int bsize, int pk,
                                           double *psi, double *h) {
    bool fw = false, bw = false, zr = false;
#pragma unroll
    for (int u = 0; u < PPL; ++u) {
        fw |= a[u] > 0.5;
        bw |= (a[u] <= 0.5) && (a[u] > 1e-15);
        zr |= a[u] <= 1e-15;
    }
    using Y = std::true_type;
    using N = std::false_type;
    double pf[PPL], hf[PPL], pb[PPL], hb[PPL], pz[PPL], hz[PPL];
    if (fw) {
        double ia[PPL], r[PPL], qv[PPL], wv[PPL];
#pragma unroll
        for (int u = 0; u < PPL; ++u) {
            const double bm = 1.0 - a[u];
            ia[u] = 1.0 / a[u];
            r[u] = -bm * ia[u];
            qv[u] = 0.0;
            wv[u] = 0.0;
            pf[u] = 0.0;
            hf[u] = 0.0;
        }
        auto step = [&](int m, auto psi_on, auto w_on) {
            const double Pm = P[m], Am = A[m], dBm = dB[m];
            double Dm = 0.0, dAm = 0.0;
            if constexpr (decltype(w_on)::value) {
                Dm = D[m];
                dAm = dA[m];
            }
#pragma unroll
            for (int u = 0; u < PPL; ++u) {
                qv[u] = Pm * ia[u] + r[u] * qv[u];
                if constexpr (decltype(psi_on)::value) pf[u] += Am * qv[u];
                hf[u] += qv[u] * dBm;
                if constexpr (decltype(w_on)::value) {
                    wv[u] = (Dm - c[u] * qv[u]) * ia[u] + r[u] * wv[u];
                    hf[u] += wv[u] * dAm;
                }
            }
        };
        int m = 0;
        for (; m < pk && m <= bsize - 2; ++m) step(m, Y{}, Y{});
        for (; m < pk; ++m) step(m, Y{}, N{});
        for (; m <= bsize - 2; ++m) step(m, N{}, Y{});
        for (; m < bsize; ++m) step(m, N{}, N{});
    }
    if (bw) {
        double ib[PPL], r[PPL], qv[PPL], wv[PPL];
        int m = bsize - 1;
        const double Ptop = P[bsize];
#pragma unroll
        for (int u = 0; u < PPL; ++u) {
            const double bm = 1.0 - a[u];
            ib[u] = 1.0 / bm;
            r[u] = -a[u] * ib[u];
            qv[u] = 0.0;
            wv[u] = 0.0;
            pb[u] = 0.0;
            hb[u] = 0.0;
            qv[u] = Ptop * ib[u];
            if (m < pk) pb[u] += A[m] * qv[u];
            hb[u] += qv[u] * dB[m];
            if (m >= 1) {
                wv[u] = (D[m] - c[u] * qv[u]) * ib[u] + r[u] * wv[u];
                hb[u] += wv[u] * dA[m - 1];
            }
        }
        auto step = [&](int m, auto psi_on, auto w_on) {
            const double Pm = P[m + 1], Am = A[m], dBm = dB[m];
            double Dm = 0.0, dAm = 0.0;
            if constexpr (decltype(w_on)::value) {
                Dm = D[m];
                dAm = dA[m - 1];
            }
#pragma unroll
            for (int u = 0; u < PPL; ++u) {
                qv[u] = Pm * ib[u] + r[u] * qv[u];
                if constexpr (decltype(psi_on)::value) pb[u] += Am * qv[u];
                hb[u] += qv[u] * dBm;
                if constexpr (decltype(w_on)::value) {
                    wv[u] = (Dm - c[u] * qv[u]) * ib[u] + r[u] * wv[u];
                    hb[u] += wv[u] * dAm;
                }
            }
        };
        --m;
        for (; m >= pk && m >= 1; --m) step(m, N{}, Y{});
        for (; m >= pk; --m) step(m, N{}, N{});
        for (; m >= 1; --m) step(m, Y{}, Y{});
        for (; m >= 0; --m) step(m, Y{}, N{});
    }
    if (zr) {
#pragma unroll
        for (int u = 0; u < PPL; ++u) {
            pz[u] = 0.0;
            hz[u] = 0.0;
        }
        for (int m = 0; m < bsize; ++m) {
            const double qv = P[m + 1], dBm = dB[m];
            const bool wo = m <= bsize - 2;
            const double Dm1 = wo ? D[m + 1] : 0.0, Pm2 = (wo && m + 2 <= B) ? P[m + 2] : 0.0;
            const double dAm = wo ? dA[m] : 0.0, Am = (m < pk) ? A[m] : 0.0;
#pragma unroll
            for (int u = 0; u < PPL; ++u) {
                if (m < pk) pz[u] += Am * qv;
                hz[u] += qv * dBm;
                if (wo) {
                    const double wv = Dm1 - c[u] * (m + 2 <= B ? Pm2 : 0.0);
                    hz[u] += wv * dAm;
                }
            }
        }
    }
#pragma unroll
    for (int u = 0; u < PPL; ++u) {
        psi[u] = a[u] > 0.5 ? pf[u] : (a[u] > 1e-15 ? pb[u] : pz[u]);
        h[u] = a[u] > 0.5 ? hf[u] : (a[u] > 1e-15 ? hb[u] : hz[u]);
    }
}

// Adjoint leaves with PPL players of one block per lane and the next
// q-plane's operands loaded while the current one is computed (as
// k_leaf_extract_n); the same results, bit for bit, as k_leaf_vjp.
template <int B, int PPL>
__global__ __launch_bounds__(256) void k_leaf_vjp_n(VjpLeafJob V) {
    const LeafJob &J = V.L;
    constexpr int LPB = B / PPL;
    constexpr int BPW = 256 / LPB;
    constexpr int NE = BPW * (B + 1);            // staged elements per array and q-plane
    constexpr int NT = (NE + 255) / 256;         // of them per thread
    static_assert(LPB * PPL == B && BPW * LPB == 256, "lanes must tile the blocks");
    __shared__ double shP[2][NE], shD[2][NE], shA[2][NE], shdA[2][NE], shdB[2][NE];
    const int t = threadIdx.x;
    const int lb = t / LPB, li = t % LPB;
    const int bfirst = blockIdx.x * BPW;
    const int b = bfirst + lb;
    const int chunk = blockIdx.y;
    const int q0 = chunk * J.q_per_chunk;
    const int q1 = min(q0 + J.q_per_chunk, J.nq);
    const int bsize = (b < J.nb) ? min(B, J.n - b * B) : 0;
    const int pk = min(J.G0, bsize);
    int jj[PPL];
    bool valid[PPL];
    double Sj[PPL], vj[PPL];
    bool any = false;
#pragma unroll
    for (int u = 0; u < PPL; ++u) {
        jj[u] = b * B + li + u * LPB;
        valid[u] = (b < J.nb) && (jj[u] < J.n);
        Sj[u] = valid[u] ? J.S[jj[u]] : 0.0;
        vj[u] = (valid[u] && V.v) ? V.v[jj[u]] : 0.0;
        any |= valid[u];
    }
    const long long bplane = (long long)J.nb * (B + 1);
    int eb[NT], em[NT];
#pragma unroll
    for (int i = 0; i < NT; ++i) {
        const int e = t + i * 256;
        eb[i] = (e < NE) ? bfirst + e / (B + 1) : J.nb;
        em[i] = e % (B + 1);
    }
    double rP[NT], rD[NT], rA[NT], rdA[NT], rdB[NT];
    auto fetch = [&](int q) {
#pragma unroll
        for (int i = 0; i < NT; ++i) {
            const bool ok = eb[i] < J.nb;
            const int m = em[i];
            const long long o = q * bplane + (long long)eb[i] * (B + 1) + m;
            rP[i] = ok ? J.blocks[o] : 0.0;
            rD[i] = ok ? V.dblocks[o] : 0.0;
            const double *ga = J.g + q * J.g_plane + (long long)eb[i] * J.g_stride;
            const double a0 = (ok && m < J.G0) ? ga[m] : 0.0;
            const double a1 = (ok && m + 1 < J.G0) ? ga[m + 1] : 0.0;
            rA[i] = a0;
            rdA[i] = a0 - a1;
            const double *gb = V.bvec ? V.bvec + q * V.b_plane + (long long)eb[i] * V.b_stride : nullptr;
            const double b0 = (ok && gb && m < V.c0) ? gb[m] : 0.0;
            const double b1 = (ok && gb && m + 1 < V.c0) ? gb[m + 1] : 0.0;
            rdB[i] = b0 - b1;
        }
    };
    auto stash = [&](int buf) {
#pragma unroll
        for (int i = 0; i < NT; ++i) {
            const int e = t + i * 256;
            if (e < NE) {
                shP[buf][e] = rP[i];
                shD[buf][e] = rD[i];
                shA[buf][e] = rA[i];
                shdA[buf][e] = rdA[i];
                shdB[buf][e] = rdB[i];
            }
        }
    };
    if (q0 < q1) {
        fetch(q0);
        stash(0);
    }
    __syncthreads();
    double accE[PPL], accD[PPL];
#pragma unroll
    for (int u = 0; u < PPL; ++u) accE[u] = accD[u] = 0.0;
    for (int q = q0; q < q1; ++q) {
        const int cur = (q - q0) & 1;
        const bool more = q + 1 < q1;
        if (more) fetch(q + 1);
        if (any) {
            const double lv = J.logv[q], wq = J.wq[q];
            const double ivv = exp(-lv);
            double a[PPL], c[PPL], pw[PPL], psi[PPL], h[PPL];
#pragma unroll
            for (int u = 0; u < PPL; ++u) {
                a[u] = !valid[u] ? 0.75 : (J.a_in ? J.a_in[(long long)q * J.n + jj[u]] : a_of(Sj[u], lv));
                pw[u] = wq * Sj[u] * a[u] * ivv;
                if (!isfinite(pw[u])) pw[u] = 0.0;
                c[u] = V.c_in ? (valid[u] ? V.c_in[(long long)q * J.n + jj[u]] : 0.0) : vj[u] * pw[u];
            }
            const int o = lb * (B + 1);
            leaf_vjp_n<B, PPL>(shP[cur] + o, shD[cur] + o, shA[cur] + o, shdA[cur] + o, shdB[cur] + o, a, c, bsize,
                               pk, psi, h);
#pragma unroll
            for (int u = 0; u < PPL; ++u) {
                if (!valid[u]) continue;
                if (J.inner_out) J.inner_out[(long long)q * J.n + jj[u]] = psi[u];
                if (V.h_out) V.h_out[(long long)q * J.n + jj[u]] = h[u];
                accE[u] += pw[u] * psi[u];
                double dterm = vj[u] * wq * ivv * a[u] * (1.0 + Sj[u] * lv) * psi[u];
                if (!isfinite(dterm)) dterm = 0.0;
                accD[u] += dterm + h[u] * a[u] * lv;
            }
        }
        if (more) stash(cur ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < PPL; ++u) {
        if (!valid[u]) continue;
        if (J.partial) J.partial[(long long)chunk * J.n + jj[u]] = accE[u];
        if (V.partial_ds) V.partial_ds[(long long)chunk * J.n + jj[u]] = accD[u];
    }
}

// Per player j of block b (A = forward g of the block, Bv = adjoint of P_b):
//   Q = P_b / f_j,  W = (dP_b - c_j Q) / f_j          (same division branch)
//   Psi = sum_{m<pk} A[m] Q[m]
//   H   = sum_m Q[m] (Bv[m] - Bv[m+1]) + sum_m W[m] (A[m] - A[m+1])   = dPhi/da_j
//   dS_j += v_j w exp(-logv) a_j (1 + S_j logv) Psi + H a_j logv
//   E_j  += w S_j a_j exp(-logv) Psi
// Q and W stream in the division's direction, so nothing is stored per lane.
template <int B>
__global__ __launch_bounds__(256) void k_leaf_vjp(VjpLeafJob V) {
    const LeafJob &J = V.L;
    constexpr int BPW = 256 / B;
    __shared__ double shP[BPW * (B + 1)];
    __shared__ double shD[BPW * (B + 1)];
    __shared__ double shA[BPW * (B + 1)];
    __shared__ double shB[BPW * (B + 2)];
    const int t = threadIdx.x;
    const int lb = t / B, li = t % B;
    const int bfirst = blockIdx.x * BPW;
    const int b = bfirst + lb;
    const int j = b * B + li;
    const int chunk = blockIdx.y;
    const int q0 = chunk * J.q_per_chunk;
    const int q1 = min(q0 + J.q_per_chunk, J.nq);
    const bool valid = (b < J.nb) && (j < J.n);
    const int bsize = (b < J.nb) ? min(B, J.n - b * B) : 0;
    const int pk = min(J.G0, bsize);
    const double Sj = valid ? J.S[j] : 0.0;
    const double vj = (valid && V.v) ? V.v[j] : 0.0;
    const long long bplane = (long long)J.nb * (B + 1);
    double accE = 0.0, accD = 0.0;
    for (int q = q0; q < q1; ++q) {
        for (int e = t; e < BPW * (B + 1); e += 256) {
            const int lbk = e / (B + 1), m = e % (B + 1);
            const int bb = bfirst + lbk;
            const bool ok = bb < J.nb;
            const long long o = q * bplane + (long long)bb * (B + 1) + m;
            shP[e] = ok ? J.blocks[o] : 0.0;
            shD[e] = ok ? V.dblocks[o] : 0.0;
            shA[e] = (ok && m < J.G0) ? J.g[q * J.g_plane + (long long)bb * J.g_stride + m] : 0.0;
        }
        for (int e = t; e < BPW * (B + 2); e += 256) {
            const int lbk = e / (B + 2), m = e % (B + 2);
            const int bb = bfirst + lbk;
            shB[e] = (bb < J.nb && m < V.c0 && V.bvec) ? V.bvec[q * V.b_plane + (long long)bb * V.b_stride + m]
                                                         : 0.0;
        }
        __syncthreads();
        if (valid) {
            const double lv = J.logv[q], wq = J.wq[q];
            const double ivv = exp(-lv);
            const double a = J.a_in ? J.a_in[(long long)q * J.n + j] : a_of(Sj, lv);
            double pw = wq * Sj * a * ivv;
            if (!isfinite(pw)) pw = 0.0;
            const double c = V.c_in ? V.c_in[(long long)q * J.n + j] : vj * pw;
            double psi, h;
            leaf_vjp_player<B>(shP + lb * (B + 1), shD + lb * (B + 1), shA + lb * (B + 1), shB + lb * (B + 2), a,
                               c, bsize, pk, psi, h);
            if (J.inner_out) J.inner_out[(long long)q * J.n + j] = psi;
            if (V.h_out) V.h_out[(long long)q * J.n + j] = h;
            accE += pw * psi;
            double dterm = vj * wq * ivv * a * (1.0 + Sj * lv) * psi;
            if (!isfinite(dterm)) dterm = 0.0;
            accD += dterm + h * a * lv;
        }
        __syncthreads();
    }
    if (valid) {
        if (J.partial) J.partial[(long long)chunk * J.n + j] = accE;
        if (V.partial_ds) V.partial_ds[(long long)chunk * J.n + j] = accD;
    }
}

// ---------------------------------------------------------------------------
// Ragged-batch adjoint leaves (global block numbering as the forward batch).
// v: cotangent in sorted concatenated order.
// ---------------------------------------------------------------------------
template <int B>
__global__ __launch_bounds__(64) void k_leaf_build_vjp_batch(const BatchInst *__restrict__ inst,
                                                             const int *__restrict__ blk_inst,
                                                             const double *__restrict__ S,
                                                             const double *__restrict__ v, double *blocks,
                                                             double *dblocks, int NB, int q0) {
    __shared__ double sh[64 * (B + 1)];
    const int lane = threadIdx.x;
    const int qi = blockIdx.y;
    const int b0 = blockIdx.x * 64;
    const int gb = b0 + lane;
    double P[B + 1], D[B + 1];
    P[0] = 1.0;
    D[0] = 0.0;
#pragma unroll
    for (int m = 1; m <= B; ++m) { P[m] = 0.0; D[m] = 0.0; }
    if (gb < NB) {
        const BatchInst &I = inst[blk_inst[gb]];
        const double lv = I.logv[q0 + qi], wq = I.w[q0 + qi];
        const double ivv = exp(-lv);
        const int b = gb - I.boff;
        // stacks and v loaded eight players at a time, unconditionally (as
        // k_leaf_build_vjp)
        constexpr int CH = 8;
        static_assert(B % CH == 0, "whole chunks");
#pragma unroll
        for (int i0 = 0; i0 < B; i0 += CH) {
            double sv[CH], vv[CH];
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const int j = b * B + i0 + u;
                const bool ok = j < I.n;
                sv[u] = *as_global(ok ? S + I.poff + j : kZeroOne);
                vv[u] = *as_global(ok ? v + I.poff + j : kZeroOne);
            }
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const int j = b * B + i0 + u;
                double a = 1.0, c = 0.0;
                if (j < I.n) {
                    const double Sj = sv[u];
                    a = a_of(Sj, lv);
                    double pw = wq * Sj * a * ivv;
                    if (!isfinite(pw)) pw = 0.0;
                    c = vv[u] * pw;
                }
                const double bb = 1.0 - a;
#pragma unroll
                for (int m = i0 + u + 1; m >= 1; --m) {
                    D[m] = a * D[m] + bb * D[m - 1] + c * P[m];
                    P[m] = a * P[m] + bb * P[m - 1];
                }
                D[0] = a * D[0] + c * P[0];
                P[0] *= a;
            }
        }
    }
    const int nblk = min(64, NB - b0);
    const long long off = (long long)qi * NB * (B + 1) + (long long)b0 * (B + 1);
#pragma unroll
    for (int m = 0; m <= B; ++m) sh[lane * (B + 1) + m] = P[m];
    __syncthreads();
    for (int e = lane; e < nblk * (B + 1); e += 64) blocks[off + e] = sh[e];
    __syncthreads();
#pragma unroll
    for (int m = 0; m <= B; ++m) sh[lane * (B + 1) + m] = D[m];
    __syncthreads();
    for (int e = lane; e < nblk * (B + 1); e += 64) dblocks[off + e] = sh[e];
}

template <int B>
__global__ __launch_bounds__(256) void k_leaf_vjp_batch(const BatchInst *__restrict__ inst,
                                                        const int *__restrict__ blk_inst,
                                                        const double *__restrict__ S,
                                                        const double *__restrict__ v,
                                                        const double *__restrict__ blocks,
                                                        const double *__restrict__ dblocks, int NB, int nq, int q0,
                                                        int q_per_chunk, double *partial, double *partial_ds,
                                                        int Ntot) {
    constexpr int BPW = 256 / B;
    __shared__ double shP[BPW * (B + 1)];
    __shared__ double shD[BPW * (B + 1)];
    __shared__ double shA[BPW * (B + 1)];
    __shared__ double shB[BPW * (B + 2)];
    const int t = threadIdx.x;
    const int lb = t / B, li = t % B;
    const int gfirst = blockIdx.x * BPW;
    const int gb = gfirst + lb;
    const int chunk = blockIdx.y;
    const int qa = chunk * q_per_chunk;
    const int qe = min(qa + q_per_chunk, nq);
    const BatchInst *I = (gb < NB) ? &inst[blk_inst[gb]] : nullptr;
    const int b = I ? gb - I->boff : 0;
    const int j = b * B + li;
    const bool valid = I && j < I->n;
    const int bsize = I ? min(B, I->n - b * B) : 0;
    const int pk = I ? min(I->G0, bsize) : 0;
    const double Sj = valid ? S[I->poff + j] : 0.0;
    const double vj = valid ? v[I->poff + j] : 0.0;
    const long long bplane = (long long)NB * (B + 1);
    double accE = 0.0, accD = 0.0;
    for (int q = qa; q < qe; ++q) {
        for (int e = t; e < BPW * (B + 1); e += 256) {
            const int g2 = gfirst + e / (B + 1), m = e % (B + 1);
            double pv = 0.0, dv = 0.0, av = 0.0;
            if (g2 < NB) {
                const long long o = q * bplane + (long long)g2 * (B + 1) + m;
                pv = blocks[o];
                dv = dblocks[o];
                const BatchInst &I2 = inst[blk_inst[g2]];
                if (m < I2.G0) av = I2.g0.base[q * I2.g0.plane + (long long)(g2 - I2.boff) * I2.g0.stride + m];
            }
            shP[e] = pv;
            shD[e] = dv;
            shA[e] = av;
        }
        for (int e = t; e < BPW * (B + 2); e += 256) {
            const int g2 = gfirst + e / (B + 2), m = e % (B + 2);
            double bv = 0.0;
            if (g2 < NB) {
                const BatchInst &I2 = inst[blk_inst[g2]];
                if (I2.b0.base && m < I2.c0)
                    bv = I2.b0.base[q * I2.b0.plane + (long long)(g2 - I2.boff) * I2.b0.stride + m];
            }
            shB[e] = bv;
        }
        __syncthreads();
        if (valid) {
            const double lv = I->logv[q0 + q], wq = I->w[q0 + q];
            const double ivv = exp(-lv);
            const double a = a_of(Sj, lv);
            double pw = wq * Sj * a * ivv;
            if (!isfinite(pw)) pw = 0.0;
            double psi, h;
            leaf_vjp_player<B>(shP + lb * (B + 1), shD + lb * (B + 1), shA + lb * (B + 1), shB + lb * (B + 2), a,
                               vj * pw, bsize, pk, psi, h);
            accE += pw * psi;
            double dterm = vj * wq * ivv * a * (1.0 + Sj * lv) * psi;
            if (!isfinite(dterm)) dterm = 0.0;
            accD += dterm + h * a * lv;
        }
        __syncthreads();
    }
    if (valid) {
        partial[(long long)chunk * Ntot + I->poff + j] = accE;
        partial_ds[(long long)chunk * Ntot + I->poff + j] = accD;
    }
}

__global__ void k_gather(const double *__restrict__ src, const int *__restrict__ perm, int n,
                         double *__restrict__ dst) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) dst[j] = src[perm[j]];
}

// Fixed-order sums over the rows of partial[rows][n]: 64 columns per
// workgroup, RG row groups (row b goes to group b mod RG, four independent
// running sums per group), groups combined in order through LDS.  The same
// association on every run: deterministic.  Optionally adds acc (the running
// total of earlier q-batches), masks inactive players, scatters to the
// original order.
template <int RG>
__global__ __launch_bounds__(64 * RG) void k_colsum(const double *__restrict__ partial, int rows, int n,
                                                    const unsigned char *__restrict__ active,
                                                    double *__restrict__ acc, int accumulate,
                                                    const int *__restrict__ perm, double *__restrict__ equity) {
    __shared__ double red[RG][64];
    const int lane = threadIdx.x & 63, grp = threadIdx.x >> 6;
    const int j = blockIdx.x * 64 + lane;
    double s = 0.0;
    if (j < n) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        int b = grp;
        for (; b + 3 * RG < rows; b += 4 * RG) {
            s0 += partial[(long long)b * n + j];
            s1 += partial[(long long)(b + RG) * n + j];
            s2 += partial[(long long)(b + 2 * RG) * n + j];
            s3 += partial[(long long)(b + 3 * RG) * n + j];
        }
        for (; b < rows; b += RG) s0 += partial[(long long)b * n + j];
        s = (s0 + s1) + (s2 + s3);
    }
    if (RG > 1) {
        red[grp][lane] = s;
        __syncthreads();
        if (grp != 0) return;
        for (int r = 1; r < RG; ++r) s += red[r][lane];
    }
    if (j >= n) return;
    if (accumulate) s += acc[j];
    if (active && !active[j]) s = 0.0;
    acc[j] = s;
    if (equity) equity[perm[j]] = s;
}

static hipError_t colsum(const double *partial, int rows, int n, const unsigned char *active, double *acc,
                         int accumulate, const int *perm, double *equity, hipStream_t s) {
    const int blocks = (n + 63) / 64;
    if (rows <= 2)
        k_colsum<1><<<blocks, 64, 0, s>>>(partial, rows, n, active, acc, accumulate, perm, equity);
    else
        k_colsum<8><<<blocks, 512, 0, s>>>(partial, rows, n, active, acc, accumulate, perm, equity);
    return hipGetLastError();
}

__global__ void k_scatter(const double *__restrict__ acc, int n, const int *__restrict__ perm,
                          double *__restrict__ equity) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    equity[perm[j]] = acc[j];
}

// ---------------------------------------------------------------------------
// Small instances: one workgroup per quadrature node, the whole tree in LDS
// (blocks of one player, schoolbook sums per level).  Level 0 is the factor
// a_j + (1 - a_j) x of each player (a_j kept at LDS offset 0); levels
// 1 .. L-2 are products truncated to c[l] coefficients; the propagate runs
// g from the payout at the root down to one entry per player, which is
// Psi_j (a level-0 node's g pairs with the empty leave-one-out product).
// The same sums as the reference's single-kernel path
// (src/gpu/gpu_kernels.cu:1547-1668) on this engine's ragged geometry, with
// the weighted contributions kept per workgroup instead of atomics.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double small_coef(const double *sm, const SmallJob &J, int l, int node, int t) {
    if (l == 0) return t == 0 ? sm[node] : 1.0 - sm[node];
    return sm[J.off[l] + node * J.c[l] + t];
}

template <int LDS>
__global__ __launch_bounds__(kSmallThreads) void k_small(SmallJob J) {
    __shared__ double sm[LDS];
    const int T = kSmallThreads, tid = threadIdx.x, n = J.n, L = J.L;
    double *row = J.partial + (long long)blockIdx.x * n;
    for (int j = tid; j < n; j += T) row[j] = 0.0;   // each j stays with one thread below
    for (int q = blockIdx.x; q < J.Qa; q += gridDim.x) {
        const double lv = J.logv[q];
        for (int j = tid; j < n; j += T) sm[j] = a_of(J.S[j], lv);
        __syncthreads();
        for (int l = 1; l <= L - 2; ++l) {           // build: parent = left * right mod x^c[l]
            const int cc = J.c[l - 1], cp = J.c[l], nrc = J.nr[l - 1], nrp = J.nr[l];
            double *P = sm + J.off[l];
            for (int e = tid; e < nrp * cp; e += T) {
                const int j = e / cp, m = e - j * cp, li = 2 * j, ri = 2 * j + 1;
                double acc = 0.0;
                if (ri >= nrc) {
                    acc = (m < cc) ? small_coef(sm, J, l - 1, li, m) : 0.0;
                } else {
                    const int t0 = max(0, m - cc + 1), t1 = min(m, cc - 1);
                    for (int t = t0; t <= t1; ++t)
                        acc += small_coef(sm, J, l - 1, li, t) * small_coef(sm, J, l - 1, ri, m - t);
                }
                P[e] = acc;
            }
            __syncthreads();
        }
        int src = -1, dst = 0;                        // g levels: -1 = the payout (root)
        for (int l = L - 1; l >= 1; --l) {            // propagate: child g = sibling (x) parent g
            const int Gp = (src < 0) ? J.k : J.g[l];
            const double *gp = (src < 0) ? J.payout : sm + J.goff[src];
            double *gc = sm + J.goff[dst];
            const int nrc = J.nr[l - 1], cc = J.c[l - 1], gcn = J.g[l - 1];
            for (int e = tid; e < nrc * gcn; e += T) {
                const int i = e / gcn, m = e - i * gcn, sib = i ^ 1;
                const double *gpar = gp + (i >> 1) * Gp;
                double acc;
                if (sib >= nrc) {
                    acc = gpar[m];                    // lone child: the missing sibling is 1
                } else {
                    acc = 0.0;
                    const int tmax = min(cc, Gp - m);
                    for (int t = 0; t < tmax; ++t) acc += small_coef(sm, J, l - 1, sib, t) * gpar[m + t];
                }
                gc[e] = acc;
            }
            __syncthreads();
            src = dst;
            dst ^= 1;
        }
        const double ivv = exp(-lv), wq = J.w[q];
        for (int j = tid; j < n; j += T) {
            const double psi = (src < 0) ? J.payout[0] : sm[J.goff[src] + j];
            double pw = wq * J.S[j] * sm[j] * ivv;    // w_q S_j a_j exp(-logv_q), non-finite -> 0
            if (!isfinite(pw)) pw = 0.0;
            row[j] += pw * psi;
        }
        __syncthreads();                              // before the next node overwrites the tree
    }
}


// Streaming probes (measured HBM rates beside the nominal peak): grid-stride
// loops over n4 double4 units; the read probe leaves one sum per workgroup so
// its loads are live.
__global__ void __launch_bounds__(256) k_stream_read(const double4 *__restrict__ src, long long n4,
                                                     double *__restrict__ part) {
    __shared__ double red[256];
    double a = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n4; i += (long long)gridDim.x * 256) {
        const double4 v = src[i];
        a += (v.x + v.y) + (v.z + v.w);
    }
    red[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// LDS read probe: each workgroup fills 32 KB of LDS from src, then every lane
// reads 16-byte entries at consecutive addresses (conflict-free ds_read_b128)
// iters times and sums every component, so each read is live.
__global__ void __launch_bounds__(256) k_lds_probe(const double2 *__restrict__ src, double *__restrict__ out,
                                                   int iters) {
    __shared__ double2 sm[2048];
    const int t = threadIdx.x;
    for (int i = t; i < 2048; i += 256) sm[i] = src[i];
    __syncthreads();
    double a = 0.0;
    for (int it = 0; it < iters; ++it) {
        const double2 v = sm[(t + it * 257) & 2047];
        a += v.x + v.y;
    }
    out[(long long)blockIdx.x * 256 + t] = a;
}

__global__ void __launch_bounds__(256) k_stream_write(double4 *__restrict__ dst, long long n4, double v) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n4; i += (long long)gridDim.x * 256)
        dst[i] = make_double4(v, v, v, v);
}

__global__ void __launch_bounds__(256) k_stream_copy(const double4 *__restrict__ src, double4 *__restrict__ dst,
                                                     long long n4) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n4; i += (long long)gridDim.x * 256)
        dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// Spectrum helpers.  A complex transform of z = x + i y (x, y real) gives
// X[K] = (Z[K] + conj(Z[-K]))/2 and Y[K] = (Z[K] - conj(Z[-K]))/(2i).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 split_x(double2 zk, double2 zm) {
    return make_double2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y));
}
__device__ __forceinline__ double2 split_y(double2 zk, double2 zm) {
    return make_double2(0.5 * (zk.y + zm.y), -0.5 * (zk.x - zm.x));
}
// A + i B for complex A, B.
__device__ __forceinline__ double2 pack_ai(double2 A, double2 Bv) {
    return make_double2(A.x - Bv.y, A.y + Bv.x);
}

// Coefficient m of node `node` in q-plane qi.
__device__ __forceinline__ double view_at(const CoefView &V, int qi, int node, int m) {
    return as_global(V.base)[(long long)qi * V.plane + (long long)node * V.stride + m];
}

// Parent handled by slot i of a propagate launch (-1: none).
__device__ __forceinline__ int slot_parent(const LevelJob &J, int i) {
    if (i >= J.np) return -1;
    return J.plist ? J.plist[i] : i;
}

// Child pair of parent p packed as L + i R (q-plane qi); a missing right
// child is the identity polynomial (or 0 with r_missing_zero); p < 0 gives
// zeros.
__device__ __forceinline__ double2 load_child_pair(const LevelJob &J, int qi, int p, int m) {
    const bool ok = p >= 0 && p < J.nr_parent;
    const bool in = ok && m < J.child.count;
    const bool has_r = (2 * p + 1) < J.nr_child;
    const CoefView &R = J.child_r.base ? J.child_r : J.child;
    const double *lx = J.child.base + (long long)qi * J.child.plane + (long long)(2 * p) * J.child.stride + m;
    const double *ry = R.base + (long long)qi * R.plane + (long long)(2 * p + 1) * R.stride + m;
    const bool one = ok && !has_r && m == 0 && !J.r_missing_zero;
    const double *px = in ? lx : kZeroOne;
    const double *py = (in && has_r) ? ry : kZeroOne + (one ? 1 : 0);
    return make_double2(*as_global(px), *as_global(py));
}

// Exact product term L[m] R[m] of parent p (m = 0 or m = c-1).
__device__ __forceinline__ double child_prod_at(const LevelJob &J, int qi, int p, int m) {
    const double l =
        as_global(J.child.base)[(long long)qi * J.child.plane + (long long)(2 * p) * J.child.stride + m];
    double r;
    if ((2 * p + 1) < J.nr_child) {
        const CoefView &R = J.child_r.base ? J.child_r : J.child;
        r = as_global(R.base)[(long long)qi * R.plane + (long long)(2 * p + 1) * R.stride + m];
    } else {
        r = (m == 0 && !J.r_missing_zero) ? 1.0 : 0.0;
    }
    return l * r;
}

__device__ __forceinline__ void put(double *dst, double v, int acc) {
    gdouble_rw *g = as_global_rw(dst);
    if (acc) *g += v;
    else *g = v;
}

// Wrapped term of a propagate output coefficient m = c-1 when F = 2c - 2:
// the cyclic correlation read Y[0] where the linear one reads Y[F].
//   side 0 (left output, correlated with R), side 1 (right output, with L).
__device__ __forceinline__ double prop_wrap_term(const LevelJob &J, int qi, int p, int side) {
    const int c = J.child.count;
    gdouble *Y = as_global(J.gpar.base + (long long)qi * J.gpar.plane + (long long)p * J.gpar.stride);
    const double yF = (J.F < J.gpar.count) ? Y[J.F] : 0.0;
    double x;
    if (side == 0) {
        if ((2 * p + 1) < J.nr_child) {
            const CoefView &R = J.child_r.base ? J.child_r : J.child;
            x = as_global(R.base)[(long long)qi * R.plane + (long long)(2 * p + 1) * R.stride + c - 1];
        } else {
            x = (c == 1 && !J.r_missing_zero) ? 1.0 : 0.0;
        }
    } else {
        const long long o = (long long)qi * J.child.plane + (long long)(2 * p) * J.child.stride + c - 1;
        x = as_global(J.child.base)[o];
    }
    return x * (yF - Y[0]);
}

// ---------------------------------------------------------------------------
// Build, LDS-resident (F <= 4096).  Item = a pair of parents (p, p+1); the
// tile holds 4096/F items.  Z_p = L + iR -> FFT -> P_p = X Y (kept in
// registers at this thread's bin pairs), same for p+1, then U = P_p + i P_p1
// -> inverse FFT -> Re = parent p, Im = parent p+1.  The transform size F may
// be one short of the linear product length (F = 2c-2): only coefficient 0
// aliases, and it is replaced by the exact L0 R0; coefficient F (if stored) is
// the exact product of the leading terms.
// ---------------------------------------------------------------------------
// LDS-level kernels: LT threads per 4096-point tile, 4096 / LT points per
// thread: 512 (radix-8 passes) or 256 (radix-16 passes, two workgroups and
// 256 VGPRs per CU).  The plain build and propagate take 256 threads where
// radix 16 needs fewer passes than radix 8 (F = 128, 256, 1024, 2048, 4096:
// one LDS round trip fewer per transform; measured 4-9% faster at
// n = 2^20), 512 at F = 512 (three passes either way; radix 16 measured 6%
// slower) and below 128; the fused adjoint kernels (which spill at 256
// VGPRs) and the ragged-batch kernels (which switch over F) run 512 threads.
constexpr int kLT = 512;
constexpr int kLP = kTile / kLT;
template <int F, int LT> struct LdsTL { using type = SwzLayout<F>; };
template <int F> struct LdsTL<F, 256> { using type = Swz16Layout<F>; };
template <int F> constexpr int lds_build_threads() { return (F >= 128 && F != 512) ? 256 : 512; }
// Largest padded tile (F = 16): LDS of the kernels that switch over F.
constexpr int kTileMaxSize = SwzLayout<16>::SIZE;

// Tile item -> (q-plane, pair of parents).
struct PlainMap {   // one q-plane per workgroup, consecutive pairs
    int pair0, qi;
    __device__ __forceinline__ void at(int item, int &q, int &pair) const {
        q = qi;
        pair = pair0 + item;
    }
};
struct FlatMap {    // v = q * npair + pair over (q-plane, pair), from v0
    int v0, npair, nv;
    __device__ __forceinline__ void at(int item, int &q, int &pair) const {
        const int v = v0 + item;
        if (v < nv) {
            q = v / npair;
            pair = v - q * npair;
        } else {
            q = 0;
            pair = 1 << 29;   // past every parent: loads give 0, nothing is stored
        }
    }
};

// The thread index for a body run inside the persistent batch kernel's item
// loop (OPQ = true): opaque to the compiler at each call, so nothing derived
// from it is hoisted out of the loop and kept live across every kind of item
// (which pushed that kernel past 128 VGPRs).  No instruction is emitted.
template <bool OPQ> __device__ __forceinline__ int body_tid() {
    int t = threadIdx.x;
    if constexpr (OPQ) {
        asm volatile("" : "+v"(t));
        __builtin_assume(t >= 0 && t < 1024);
    }
    return t;
}

template <int F, class Map, int LT = kLT, bool OPQ = false>
__device__ __forceinline__ void build_lds_body(const LevelJob &J, double2 *lds, const Map &M, const double2 *tw,
                                               int twShift) {
    using TL = typename LdsTL<F, LT>::type;
    constexpr int LP = kTile / LT;
    constexpr int HALF = F / 2;
    constexpr int LOGH = ilog2c(HALF);
    const int t = body_tid<OPQ>();
    (void)twShift;

    double2 P1[LP / 2];
#pragma unroll
    for (int which = 0; which < 2; ++which) {
        // first pass straight from the children (global -> registers)
        auto ld = [&](int item, int m) {
            int qi, pr;
            M.at(item, qi, pr);
            return load_child_pair(J, qi, 2 * pr + which, m);
        };
        fft_tile_io<F, false, LT, TL>(lds, tw, ld, NoIO{});
#pragma unroll
        for (int u = 0; u < LP / 2; ++u) {
            const int ps = t + u * LT;
            const int item = ps >> LOGH, jb = ps & (HALF - 1);
            const int ia = TL::at(item, jb), ib = TL::at(item, jb == 0 ? HALF : F - jb);
            const double2 za = lds[ia], zb = lds[ib];
            double2 Pv;
            if (jb == 0) Pv = make_double2(za.x * za.y, zb.x * zb.y);
            else Pv = c_mul(split_x(za, zb), split_y(za, zb));
            if (which == 0) {
                P1[u] = Pv;
            } else if (jb == 0) {
                lds[ia] = make_double2(P1[u].x, Pv.x);
                lds[ib] = make_double2(P1[u].y, Pv.y);
            } else {
                lds[ia] = pack_ai(P1[u], Pv);
                lds[ib] = pack_ai(c_conj(P1[u]), c_conj(Pv));
            }
        }
    }
    // last pass straight to the parents (exact coefficient 0 and top term)
    const double invF = 1.0 / (double)F;
    const int ps_ = J.parent.stride, pc = J.parent.count;
    auto st = [&](int item, int m, double2 v) {
        int qi, pr;
        M.at(item, qi, pr);
        const int p = 2 * pr;
        double *pb = J.parent.base + (long long)qi * J.parent.plane;
        if (m < pc) {
            if (p < J.nr_parent)
                put(&pb[(long long)p * ps_ + m], (m == 0) ? child_prod_at(J, qi, p, 0) : v.x * invF, J.acc_out);
            if (p + 1 < J.nr_parent)
                put(&pb[(long long)(p + 1) * ps_ + m], (m == 0) ? child_prod_at(J, qi, p + 1, 0) : v.y * invF,
                    J.acc_out);
        }
        if (m == 0 && pc > F) {
            const int top = J.child.count - 1;
            if (p < J.nr_parent) put(&pb[(long long)p * ps_ + F], child_prod_at(J, qi, p, top), J.acc_out);
            if (p + 1 < J.nr_parent)
                put(&pb[(long long)(p + 1) * ps_ + F], child_prod_at(J, qi, p + 1, top), J.acc_out);
        }
    };
    fft_tile_io<F, true, LT, TL>(lds, tw, NoIO{}, st);
}

// Two workgroups per CU (64 KiB of LDS each): 128 VGPRs at 512 threads,
// 256 at 256 threads.
template <int F>
__global__ __launch_bounds__(lds_build_threads<F>(), lds_build_threads<F>() == 256 ? 2 : 4) void k_build_lds(
    LevelJob J, const double2 *tw, int twShift) {
    __shared__ double2 lds[kTile];
    build_lds_body<F, PlainMap, lds_build_threads<F>()>(
        J, lds, PlainMap{(int)blockIdx.x * (kTile / F), (int)blockIdx.y}, tw, twShift);
}

// ---------------------------------------------------------------------------
// Propagate, LDS-resident (F <= 4096): for parent p with children L, R,
//   gL[m] = sum_t R[t] g[m+t],  gR[m] = sum_t L[t] g[m+t],  m < G_child,
// as cyclic correlations of size F (F >= G_child + c_child - 1, so no index
// wraps).  Item = a pair of parents: their g vectors share one transform;
// each parent's children share another; the two outputs of a parent share
// the inverse (Re = gL, Im = gR).
// ---------------------------------------------------------------------------
// The g and child transforms take their first pass straight from global
// memory and the inverses store straight to the children's g (two LDS round
// trips fewer per transform); with per-pass LDS addresses this fits 128
// VGPRs at 512 threads (two workgroups per CU) without spills.
//
// Slot parents: without a parent list (PL = false) slot i is parent i, found
// by arithmetic; with one (PL = true) the launch fills spar[2 item + w] with
// the parent of the item's slot w before the body runs, so the per-point
// loaders read LDS: a list entry loaded from global memory inside a loader
// made every point wait for all earlier loads (in-order vector memory
// counters), one memory latency per point.
template <int F, class Map, bool OPQ = false, bool PL = false, int LT = kLT>
__device__ __forceinline__ void prop_lds_body(const LevelJob &J, double2 *lds, const Map &M, const double2 *tw,
                                              int twShift, const int *spar = nullptr) {
    using TL = typename LdsTL<F, LT>::type;
    constexpr int LP = kTile / LT;
    constexpr int HALF = F / 2;
    constexpr int LOGH = ilog2c(HALF);
    const int t = body_tid<OPQ>();
    const int gs = J.gpar.stride, gcnt = J.gpar.count;
    auto parent = [&](int item, int pr, int w) -> int {
        if constexpr (PL) {
            (void)pr;
            return spar[2 * item + w];
        } else {
            (void)item;
            const int i = 2 * pr + w;
            return i < J.np ? i : -1;
        }
    };

    auto ldg = [&](int item, int m) {
        int qi, pr;
        M.at(item, qi, pr);
        gdouble *gb = as_global(J.gpar.base + (long long)qi * J.gpar.plane);
        const int pa = parent(item, pr, 0), pb = parent(item, pr, 1);
        double2 z = make_double2(0.0, 0.0);
        if (m < gcnt) {
            if (pa >= 0) z.x = gb[(long long)pa * gs + m];
            if (pb >= 0) z.y = gb[(long long)pb * gs + m];
        }
        return z;
    };
    fft_tile_io<F, false, LT, TL>(lds, tw, ldg, NoIO{});
    double2 G1[LP / 2], G2[LP / 2];
#pragma unroll
    for (int u = 0; u < LP / 2; ++u) {
        const int ps = t + u * LT;
        const int item = ps >> LOGH, jb = ps & (HALF - 1);
        const double2 za = lds[TL::at(item, jb)], zb = lds[TL::at(item, jb == 0 ? HALF : F - jb)];
        if (jb == 0) {
            G1[u] = make_double2(za.x, zb.x);
            G2[u] = make_double2(za.y, zb.y);
        } else {
            G1[u] = split_x(za, zb);
            G2[u] = split_y(za, zb);
        }
    }
    const double invF = 1.0 / (double)F;
    const int gcs = J.gchild.stride, gcc = J.gchild.count;
#pragma unroll
    for (int which = 0; which < 2; ++which) {
        auto ldc = [&](int item, int m) {
            int qi, pr;
            M.at(item, qi, pr);
            return load_child_pair(J, qi, parent(item, pr, which), m);
        };
        fft_tile_io<F, false, LT, TL>(lds, tw, ldc, NoIO{});
#pragma unroll
        for (int u = 0; u < LP / 2; ++u) {
            const int ps = t + u * LT;
            const int item = ps >> LOGH, jb = ps & (HALF - 1);
            const double2 g = which == 0 ? G1[u] : G2[u];
            const int ia = TL::at(item, jb), ib = TL::at(item, jb == 0 ? HALF : F - jb);
            const double2 za = lds[ia], zb = lds[ib];
            if (jb == 0) {
                lds[ia] = make_double2(g.x * za.y, g.x * za.x);
                lds[ib] = make_double2(g.y * zb.y, g.y * zb.x);
            } else {
                const double2 Lh = split_x(za, zb), Rh = split_y(za, zb);
                const double2 A = c_mulc(g, Rh), Bv = c_mulc(g, Lh);
                lds[ia] = pack_ai(A, Bv);
                lds[ib] = pack_ai(c_conj(A), c_conj(Bv));
            }
        }
        auto st = [&](int item, int m, double2 v) {
            int qi, pr;
            M.at(item, qi, pr);
            const int p = parent(item, pr, which);
            if (p >= 0 && m < gcc) {
                double *gcb = J.gchild.base + (long long)qi * J.gchild.plane;
                double vl = v.x * invF, vr = v.y * invF;
                if (J.gfix && m == J.child.count - 1) {
                    vl += prop_wrap_term(J, qi, p, 0);
                    vr += prop_wrap_term(J, qi, p, 1);
                }
                put(&gcb[(long long)(2 * p) * gcs + m], vl, J.acc_out);
                if (2 * p + 1 < J.nr_child) put(&gcb[(long long)(2 * p + 1) * gcs + m], vr, J.acc_out);
            }
        };
        fft_tile_io<F, true, LT, TL>(lds, tw, NoIO{}, st);
    }
}

template <int F> constexpr int lds_prop_threads() { return lds_build_threads<F>(); }
template <int F, bool PL>
__global__ __launch_bounds__(lds_prop_threads<F>(), lds_prop_threads<F>() == 256 ? 2 : 4) void k_prop_lds(
    LevelJob J, const double2 *tw, int twShift) {
    constexpr int ITEMS = kTile / F, LT = lds_prop_threads<F>();
    __shared__ double2 lds[kTile];
    __shared__ int spar[PL ? 2 * ITEMS : 1];
    const PlainMap M{(int)blockIdx.x * ITEMS, (int)blockIdx.y};
    if constexpr (PL) {
        for (int e = threadIdx.x; e < 2 * ITEMS; e += LT) {
            int qi, pr;
            M.at(e >> 1, qi, pr);
            spar[e] = slot_parent(J, 2 * pr + (e & 1));
        }
        __syncthreads();
    }
    prop_lds_body<F, PlainMap, false, PL, LT>(J, lds, M, tw, twShift, spar);
}

// ---------------------------------------------------------------------------
// Schoolbook levels (few coefficients per node).  Direct sums, no transform:
//   build      P[m]  = sum_t L[t] R[m-t],           m < c_parent
//   propagate  gL[m] = sum_t R[t] g[m+t], gR[m] = sum_t L[t] g[m+t],  m < G_child
// A workgroup of 256 threads stages the operands of ppw parents (tile items
// given by the same maps as the transform kernels, "pair" = parent) in LDS;
// a thread owns 4 consecutive outputs of one parent and slides a 4-wide
// register window over R (build) or g (propagate), so each LDS load feeds
// 4 (build) or 8 (propagate) FMAs and the broadcast coefficient is shared by
// the lanes of a parent.  Zero padding in LDS replaces bounds tests in the
// inner loop.  A missing right child is the identity polynomial.  NT threads
// share a workgroup's parents (256 in the schoolbook kernels, 512 in the
// persistent batch kernel); each output's sum is the same either way.
// ---------------------------------------------------------------------------
constexpr int kSbT = 256;

template <class Map, int NT = kSbT, bool OPQ = false>
__device__ __forceinline__ void build_sb_body(const LevelJob &J, double *sh, const Map &M, int ppw) {
    const int t = body_tid<OPQ>();
    const int cc = J.child.count, pc = J.parent.count;
    const int stride = sb_stride_build(cc);     // L[cc] | 4 zeros, R[cc], 6 zeros
    for (int e = t; e < ppw * stride; e += NT) {
        const int sl = e / stride, x = e - sl * stride;
        int qi, p;
        M.at(sl, qi, p);
        double v = 0.0;
        if (p < J.nr_parent) {
            const bool has_r = (2 * p + 1) < J.nr_child;
            if (x < cc) {
                v = view_at(J.child, qi, 2 * p, x);
            } else {
                const int r = x - cc - 4;          // R index
                if (r >= 0 && r < cc) v = has_r ? view_at(J.child, qi, 2 * p + 1, r) : (r == 0 ? 1.0 : 0.0);
            }
        }
        sh[e] = v;
    }
    __syncthreads();
    const int nqd = (pc + 3) >> 2;
    for (int task = t; task < ppw * nqd; task += NT) {
        const int sl = task / nqd, m0 = (task - sl * nqd) * 4;
        int qi, p;
        M.at(sl, qi, p);
        if (p >= J.nr_parent) continue;
        const double *L = sh + sl * stride;
        const double *Rz = L + cc + 4;             // Rz[x] = R[x], zero for x in [-4, -1] and [cc, cc+5]
        const int tlo = max(0, m0 - cc + 1), thi = min(cc - 1, m0 + 3);
        double r0 = Rz[m0 - tlo], r1 = Rz[m0 + 1 - tlo], r2 = Rz[m0 + 2 - tlo], r3 = Rz[m0 + 3 - tlo];
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        for (int tt = tlo; tt <= thi; ++tt) {
            const double l = L[tt];
            a0 = fma(l, r0, a0);
            a1 = fma(l, r1, a1);
            a2 = fma(l, r2, a2);
            a3 = fma(l, r3, a3);
            r3 = r2;
            r2 = r1;
            r1 = r0;
            r0 = Rz[m0 - tt - 1];
        }
        double *out = J.parent.base + (long long)qi * J.parent.plane + (long long)p * J.parent.stride;
        if (m0 < pc) put(&out[m0], a0, J.acc_out);
        if (m0 + 1 < pc) put(&out[m0 + 1], a1, J.acc_out);
        if (m0 + 2 < pc) put(&out[m0 + 2], a2, J.acc_out);
        if (m0 + 3 < pc) put(&out[m0 + 3], a3, J.acc_out);
    }
}

template <class Map, int NT = kSbT, bool OPQ = false>
__device__ __forceinline__ void prop_sb_body(const LevelJob &J, double *sh, const Map &M, int ppw) {
    const int t = body_tid<OPQ>();
    const int cc = J.child.count, gc = J.gchild.count, gp = J.gpar.count;
    const int glen = gc + cc + 8;
    const int stride = sb_stride_prop(cc, gc);  // L[cc], R[cc], g[gc + cc + 8]
    for (int e = t; e < ppw * stride; e += NT) {
        const int sl = e / stride, x = e - sl * stride;
        int qi, pr;
        M.at(sl, qi, pr);
        const int p = slot_parent(J, pr);
        double v = 0.0;
        if (p >= 0) {
            const bool has_r = (2 * p + 1) < J.nr_child;
            if (x < cc) {
                v = view_at(J.child, qi, 2 * p, x);
            } else if (x < 2 * cc) {
                const int r = x - cc;
                v = has_r ? view_at(J.child, qi, 2 * p + 1, r) : (r == 0 ? 1.0 : 0.0);
            } else {
                const int gx = x - 2 * cc;
                if (gx < gp && gx < glen) v = view_at(J.gpar, qi, p, gx);
            }
        }
        sh[e] = v;
    }
    __syncthreads();
    const int nqd = (gc + 3) >> 2;
    for (int task = t; task < ppw * nqd; task += NT) {
        const int sl = task / nqd, m0 = (task - sl * nqd) * 4;
        int qi, pr;
        M.at(sl, qi, pr);
        const int p = slot_parent(J, pr);
        if (p < 0) continue;
        const double *L = sh + sl * stride;
        const double *R = L + cc;
        const double *g = R + cc;
        double w0 = g[m0], w1 = g[m0 + 1], w2 = g[m0 + 2], w3 = g[m0 + 3];
        double l0 = 0.0, l1 = 0.0, l2 = 0.0, l3 = 0.0, q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0;
        for (int tt = 0; tt < cc; ++tt) {
            const double rr = R[tt], ll = L[tt];
            l0 = fma(rr, w0, l0);
            l1 = fma(rr, w1, l1);
            l2 = fma(rr, w2, l2);
            l3 = fma(rr, w3, l3);
            q0 = fma(ll, w0, q0);
            q1 = fma(ll, w1, q1);
            q2 = fma(ll, w2, q2);
            q3 = fma(ll, w3, q3);
            w0 = w1;
            w1 = w2;
            w2 = w3;
            w3 = g[m0 + 4 + tt];
        }
        double *gl = J.gchild.base + (long long)qi * J.gchild.plane + (long long)(2 * p) * J.gchild.stride;
        const bool has_r = (2 * p + 1) < J.nr_child;
        double *gr = gl + J.gchild.stride;
        const double vl[4] = {l0, l1, l2, l3}, vr[4] = {q0, q1, q2, q3};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (m0 + i < gc) {
                put(&gl[m0 + i], vl[i], J.acc_out);
                if (has_r) put(&gr[m0 + i], vr[i], J.acc_out);
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_build_sb(LevelJob J, int ppw) {
    __shared__ double sh[kSbLds];
    build_sb_body(J, sh, PlainMap{(int)blockIdx.x * ppw, (int)blockIdx.y}, ppw);
}

__global__ __launch_bounds__(256) void k_prop_sb(LevelJob J, int ppw) {
    __shared__ double sh[kSbLds];
    prop_sb_body(J, sh, PlainMap{(int)blockIdx.x * ppw, (int)blockIdx.y}, ppw);
}

__global__ __launch_bounds__(256) void k_build_sb_batch(const LevelJob *__restrict__ jobs,
                                                        const int2 *__restrict__ items) {
    __shared__ double sh[kSbLds];
    const int2 it = items[blockIdx.x];
    const LevelJob J = jobs[it.x];
    const int ppw = sb_ppw(sb_stride_build(J.child.count));
    build_sb_body(J, sh, FlatMap{it.y, J.nr_parent, J.nq * J.nr_parent}, ppw);
}

__global__ __launch_bounds__(256) void k_prop_sb_batch(const LevelJob *__restrict__ jobs,
                                                       const int2 *__restrict__ items) {
    __shared__ double sh[kSbLds];
    const int2 it = items[blockIdx.x];
    const LevelJob J = jobs[it.x];
    const int ppw = sb_ppw(sb_stride_prop(J.child.count, J.gchild.count));
    prop_sb_body(J, sh, FlatMap{it.y, J.np, J.nq * J.np}, ppw);
}

// ---------------------------------------------------------------------------
// Ragged batches: one launch per tree level over every instance that has the
// level.  Workgroup w takes work item items[w] = (job, v0): instance job's
// level, tile items v0 .. v0 + 4096/F - 1 over its (q-plane, pair) space, so
// small instances fill tiles with q-planes instead of leaving them empty.
// F is uniform per workgroup (one job), so the switch does not diverge.
// ---------------------------------------------------------------------------
#define ICM_F_CASES(BODY)                                                     \
    switch (J.F) {                                                            \
    case 16: BODY(16); break;                                                 \
    case 32: BODY(32); break;                                                 \
    case 64: BODY(64); break;                                                 \
    case 128: BODY(128); break;                                               \
    case 256: BODY(256); break;                                               \
    case 512: BODY(512); break;                                               \
    case 1024: BODY(1024); break;                                             \
    case 2048: BODY(2048); break;                                             \
    case 4096: BODY(4096); break;                                             \
    default: break;                                                           \
    }

__global__ __launch_bounds__(512) void k_build_lds_batch(const LevelJob *__restrict__ jobs,
                                                         const int2 *__restrict__ items, const double2 *tw,
                                                         int twShift) {
    __shared__ double2 lds[kTileMaxSize];
    const int2 it = items[blockIdx.x];
    const LevelJob J = jobs[it.x];
    const int npair = (J.nr_parent + 1) >> 1;
    const FlatMap M{it.y, npair, J.nq * npair};
#define ICM_BUILD_F(FF) build_lds_body<FF>(J, lds, M, tw, twShift)
    ICM_F_CASES(ICM_BUILD_F)
#undef ICM_BUILD_F
}

__global__ __launch_bounds__(512) void k_prop_lds_batch(const LevelJob *__restrict__ jobs,
                                                        const int2 *__restrict__ items, const double2 *tw,
                                                        int twShift) {
    __shared__ double2 lds[kTileMaxSize];
    const int2 it = items[blockIdx.x];
    const LevelJob J = jobs[it.x];
    const int npair = (J.np + 1) >> 1;
    const FlatMap M{it.y, npair, J.nq * npair};
#define ICM_PROP_F(FF) prop_lds_body<FF>(J, lds, M, tw, twShift)
    ICM_F_CASES(ICM_PROP_F)
#undef ICM_PROP_F
}

// ---------------------------------------------------------------------------
// Persistent ragged batch: one launch runs a q-pass's leaf build, every tree
// level, the leaf extract and the fixed-order q reduction of all instances.
//
// XCD-local queues.  MI355X has eight XCDs, each with its own L2; a write is
// seen by another XCD only after an L2 write-back there and an invalidate
// here, which per work item cost 3.4x the whole batch (measured).  So the
// host splits the instances into eight contiguous ranges of about equal
// work, one item queue per XCD, and a workgroup serves the queue of the XCD
// it runs on (hardware XCC id).  An instance's whole tree then lives in one
// L2, and a dependency between items needs no L2 write-back: the producer
// waits until its stores are complete (in L2) before the completion count
// is raised, and the consumer drops its CU's L1 lines before reading.  The
// L1 drop takes the device-scope cache invalidate (L1 and the L2's
// non-coherent lines): the workgroup-scope one left stale L1 lines of reused
// regions in place (wrong results in runs with small q-passes, where a
// region fits a CU's L1 between its reuses).  Counters and tickets are
// device-scope atomics.
//
// Within a queue, workgroups take items by ticket.  Items are ordered phase
// by phase (leaves; build levels 1, 2, ..; propagate levels .., 2, 1;
// extract; finalize), so every item's inputs come from items with smaller
// tickets; ctr[i] counts the finished items of instance i, and an item starts
// once its instance (every instance it spans) has finished the items of all
// its earlier phases (`need`): an instance's next level starts as soon as its
// own previous level is done, with no barrier between levels.  Progress: a
// workgroup takes a ticket only while it runs, so the unfinished item with
// the smallest ticket of a queue always has its inputs and proceeds.  Every
// wait is further bounded by spin_limit (past it the error flag is set, the
// item runs anyway and the host reports the error); a queue that no
// workgroup served (an XCD without workgroups) is reported the same way by
// the last workgroup to finish.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int xcc_id() {
    int x;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID, 0, 4)" : "=s"(x));
    return x;
}

// This wave's global stores are complete (in its XCD's L2).
__device__ __forceinline__ void stores_complete() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }
// Drop this CU's L1 lines (and the L2's non-coherent ones): later loads of
// any wave on the CU read the XCD's L2.  Completed before returning.
__device__ __forceinline__ void cache_invalidate() { asm volatile("buffer_inv sc1\n\ts_waitcnt vmcnt(0)" ::: "memory"); }

__device__ __forceinline__ double dpp_row_shr1(double v) {
    // lane l receives lane l - 1's value within each row of 16 lanes (row
    // lane 0 receives 0); used within groups of 8 lanes, whose lane 0 ignores it
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(x & 0xffffffffLL), 0x111, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(x >> 32), 0x111, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Leaf build of the 64 global blocks gb0 .. gb0 + 63 (those below gend) for
// q-planes qa .. qe - 1 of the pass, eight lanes per block (512 threads): lane r holds coefficients
// r CPL .. r CPL + CPL - 1 (lane 7 also the top one, B).  Player i multiplies
// by a + (1 - a) x, P[m] = a P[m] + (1 - a) P[m - 1], with the same operations
// as the one-lane recurrence of k_leaf_build_batch (coefficients above the
// current degree stay exactly 0); the coefficient below a lane's first comes
// from the lane before it.  128 VGPRs suffice where one lane per block needs
// B + 1 coefficients in registers.  The a values go through LDS (each
// computed once, double-buffered by q-plane, pairs XOR-swizzled by block so
// the eight blocks of a wave read distinct banks).
template <int B>
__device__ __forceinline__ void leaf_build_coop(const BatchInst *__restrict__ inst, const int *__restrict__ blk_inst,
                                                const double *__restrict__ S, double *__restrict__ blocks, int NB,
                                                int gend, int qa, int qe, int q0, int gb0, double *sa) {
    constexpr int CPL = B / 8;
    static_assert(CPL * 8 == B && 2 * 64 * B <= 2 * kTileMaxSize, "a values of two q-planes must fit the LDS");
    static_assert(CPL + 1 <= 16, "a lane's stores must fit kSink");
    const int t = body_tid<true>(), bl = t >> 3, r = t & 7, sw = bl & 7;
    const int gb = gb0 + bl;
    const bool has = gb < gend;
    const BatchInst *I = has ? &inst[blk_inst[gb]] : nullptr;
    const int b = has ? gb - I->boff : 0;
    double Sv[CPL];
    bool pv[CPL];
#pragma unroll
    for (int u = 0; u < CPL; ++u) {
        const int j = b * B + r * CPL + u;
        pv[u] = has && j < I->n;
        Sv[u] = *as_global(pv[u] ? S + I->poff + j : kZeroOne);
    }
    const long long bplane = (long long)NB * (B + 1);
    // the next plane's node is read before this plane's stores: read after
    // them it waited for them (stores count in the same in-order counter)
    gdouble *nodes = as_global(has ? I->logv + q0 : kZeroOne);
    const int qmul = has ? 1 : 0;
    double lv = (qa < qe) ? nodes[qa * qmul] : 0.0;
    vm_drain();
    for (int qi = qa; qi < qe; ++qi) {
        double *sq = sa + (qi & 1) * (64 * B) + bl * B;
#pragma unroll
        for (int u = 0; u < CPL; ++u) {
            const int i = r * CPL + u;
            sq[2 * ((i >> 1) ^ sw) + (i & 1)] = pv[u] ? a_of(Sv[u], lv) : 1.0;
        }
        __syncthreads();
        lv = nodes[min(qi + 1, qe - 1) * qmul];
        double P[CPL], top = 0.0;
#pragma unroll
        for (int u = 0; u < CPL; ++u) P[u] = 0.0;
        if (r == 0) P[0] = 1.0;
#pragma unroll 2
        for (int i2 = 0; i2 < B / 2; ++i2) {
            const double2 ap = *reinterpret_cast<const double2 *>(sq + 2 * (i2 ^ sw));
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const double a = h ? ap.y : ap.x;
                const double bb = 1.0 - a;
                const double left = dpp_row_shr1(P[CPL - 1]);
                top = a * top + bb * P[CPL - 1];
#pragma unroll
                for (int u = CPL - 1; u >= 1; --u) P[u] = a * P[u] + bb * P[u - 1];
                P[0] = (r == 0) ? P[0] * a : a * P[0] + bb * left;
            }
        }
        // every lane stores CPL + 1 values (lanes without a block, and the
        // top coefficient of lanes other than the last, into kSink): with
        // a lane-dependent store count the next node's read, issued before
        // these stores, waited for them
        double *dst = has ? blocks + qi * bplane + (long long)gb * (B + 1) + r * CPL : kSink;
#pragma unroll
        for (int u = 0; u < CPL; ++u) dst[u] = P[u];
        *((has && r == 7) ? dst + CPL : kSink + CPL) = top;
    }
}

// Fixed-order q reduction of one player (the association of k_colsum: row
// groups b mod RG, four running sums per group, groups added in order), then
// the running total of earlier passes, and on the last pass the scatter to
// the original order.
__device__ __forceinline__ double colsum_group(const double *__restrict__ partial, int rows, long long n, int j,
                                               int grp, int RG) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int b = grp;
    for (; b + 3 * RG < rows; b += 4 * RG) {
        s0 += partial[(long long)b * n + j];
        s1 += partial[(long long)(b + RG) * n + j];
        s2 += partial[(long long)(b + 2 * RG) * n + j];
        s3 += partial[(long long)(b + 3 * RG) * n + j];
    }
    for (; b < rows; b += RG) s0 += partial[(long long)b * n + j];
    return (s0 + s1) + (s2 + s3);
}

__device__ __forceinline__ void persist_final(const PersistArgs &A, int j0, int jend) {
    const int j = j0 + (int)threadIdx.x;
    if (j >= jend) return;
    const int rows = A.nchunk;
    double s;
    if (rows <= 2) {
        s = colsum_group(A.partial, rows, A.Ntot, j, 0, 1);
    } else {
        s = colsum_group(A.partial, rows, A.Ntot, j, 0, 8);
        for (int g = 1; g < 8; ++g) s += colsum_group(A.partial, rows, A.Ntot, j, g, 8);
    }
    if (A.accumulate) s += A.acc[j];
    A.acc[j] = s;
    if (A.equity) A.equity[A.perm[j]] = s;
}

// Batch leaf build with eight lanes per block (the per-level launches' leaf
// phase at B = 64): 64 blocks per workgroup, q-planes blockIdx.y * qper ..
// + qper - 1.  The same products as k_leaf_build_batch.
template <int B>
__global__ __launch_bounds__(kLT) void k_leaf_build_batch_coop(const BatchInst *__restrict__ inst,
                                                               const int *__restrict__ blk_inst,
                                                               const double *__restrict__ S, double *blocks, int NB,
                                                               int nq, int q0, int qper) {
    __shared__ double sa[2 * 64 * B];
    const int qa = blockIdx.y * qper;
    leaf_build_coop<B>(inst, blk_inst, S, blocks, NB, NB, qa, min(nq, qa + qper), q0, blockIdx.x * 64, sa);
}

// Waits until *ctr >= need.  Past `limit` polls it sets the error flag and
// returns; once the flag is set (by any workgroup) waits return at once, so a
// broken run drains after one bounded wait instead of one per item.
__device__ __forceinline__ void persist_wait(int *ctr, int need, int *err, int limit) {
    int spins = 0;
    while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
        if (++spins > limit || ((spins & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED,
                                                                        __HIP_MEMORY_SCOPE_AGENT))) {
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
        __builtin_amdgcn_s_sleep(2);
    }
}

constexpr int kPersistLds = kTileMaxSize * (int)sizeof(double2);

template <int B> struct PersistCfg {
    static constexpr int PPL = (B == 64) ? 2 : 1;             // players per lane of the extract
    using X = LeafExN<B, PPL>;
    static constexpr int EXB = 2 * X::BPW;                     // blocks per extract item (two halves)
    static_assert(2 * 2 * (X::NP + X::NG) <= 2 * kTileMaxSize, "extract halves must fit the LDS");
};

// The persistent kernel's LDS: dynamic, one 64 KB array that every item body
// addresses by name (LDS instructions also in the functions kept out of line).
extern __shared__ double2 persist_lds[];

// The leaf build and the extract as functions of their own (not inlined):
// inlined beside the level bodies they pushed the kernel's register
// allocation into hundreds of spills, although each fits 128 VGPRs alone.
// A few thousand such items per pass make the call cost negligible.
// Arguments arrive in VGPRs: the loop bounds are made wave-uniform again.
template <int B>
__device__ __attribute__((noinline)) void persist_leaf(const BatchInst *__restrict__ inst,
                                                       const int *__restrict__ blk_inst,
                                                       const double *__restrict__ S, double *__restrict__ blocks,
                                                       int NB, int gend, int nq, int q0, int gb0) {
    NB = __builtin_amdgcn_readfirstlane(NB);
    gend = __builtin_amdgcn_readfirstlane(gend);
    nq = __builtin_amdgcn_readfirstlane(nq);
    q0 = __builtin_amdgcn_readfirstlane(q0);
    gb0 = __builtin_amdgcn_readfirstlane(gb0);
    leaf_build_coop<B>(inst, blk_inst, S, blocks, NB, gend, 0, nq, q0, gb0, reinterpret_cast<double *>(persist_lds));
}

template <int B>
__device__ __attribute__((noinline)) void persist_extract(const BatchInst *__restrict__ inst,
                                                          const int *__restrict__ blk_inst,
                                                          const double *__restrict__ S,
                                                          const double *__restrict__ blocks, int NB, int gend, int nq,
                                                          int q0, int q_per_chunk, double *partial, int Ntot, int gb0,
                                                          int chunk) {
    using C = PersistCfg<B>;
    using X = typename C::X;
    NB = __builtin_amdgcn_readfirstlane(NB);
    gend = __builtin_amdgcn_readfirstlane(gend);
    nq = __builtin_amdgcn_readfirstlane(nq);
    q0 = __builtin_amdgcn_readfirstlane(q0);
    q_per_chunk = __builtin_amdgcn_readfirstlane(q_per_chunk);
    Ntot = __builtin_amdgcn_readfirstlane(Ntot);
    gb0 = __builtin_amdgcn_readfirstlane(gb0);
    chunk = __builtin_amdgcn_readfirstlane(chunk);
    const int t = threadIdx.x;
    double *base = reinterpret_cast<double *>(persist_lds) + (t >> 8) * 2 * (X::NP + X::NG);
    leaf_extract_batch_n_body<B, C::PPL>(t & 255, gb0 + (t >> 8) * X::BPW, chunk, inst, blk_inst, S, blocks, NB,
                                         gend, nq, q0, q_per_chunk, partial, Ntot,
                                         reinterpret_cast<double(*)[X::NP]>(base),
                                         reinterpret_cast<double(*)[X::NG]>(base + 2 * X::NP));
}

// Dynamic LDS: kPersistLds bytes (the 4096-point tile, 64 KB).  128 VGPRs
// (four waves per SIMD): two workgroups of 512 threads per CU, as the
// per-level batch kernels.
template <int B>
__global__ __launch_bounds__(kLT, 4) void k_batch_persist(PersistArgs A) {
    double2 *lds = persist_lds;
    __shared__ int s_tk, s_last;
    // No branch of the item loop may be loop-invariant and divergent: the
    // compiler would unswitch the loop on it, and a wave runs the two loop
    // copies one after the other -- lanes without the ticket fetch (t != 0)
    // would cycle through barriers and items on a ticket that never
    // advances while lane 0 waits for them.  So the thread index is taken
    // afresh in every iteration (opaque), and the ticket and item fields are
    // made wave-uniform with readfirstlane.
    // An item is counted done at the top of the next iteration, after the
    // next ticket is taken: its stores drain meanwhile (and before any wait
    // of the next item, which may need it).
    int pend = -1, pend_i1 = 0;   // kind_inst and last instance of the item to count
    for (;;) {
        int t = threadIdx.x;
        asm volatile("" : "+v"(t));
        // the queue of this workgroup's XCD (read afresh: fewer values live across the items)
        const int x = __builtin_amdgcn_readfirstlane(xcc_id()) % A.nqueues;
        if (t == 0) s_tk = atomicAdd(&A.sched[x], 1);
        __syncthreads();
        const int tk = __builtin_amdgcn_readfirstlane(s_tk);
        __syncthreads();
        if (pend >= 0) {   // outputs complete in this XCD's L2, then count the item done
            stores_complete();
            __syncthreads();
            const int pk = pend & 7, p0 = pend >> 3;
            if (pk == kPLeaf || pk == kPExtract) {
                if (p0 + t <= pend_i1)
                    __hip_atomic_fetch_add(&A.ctr[p0 + t], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else if (t == 0) {
                __hip_atomic_fetch_add(&A.ctr[p0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            pend = -1;
        }
        if (tk >= A.qoff[x + 1] - A.qoff[x]) break;
        PItem it = A.items[A.qoff[x] + tk];
        it.kind_inst = __builtin_amdgcn_readfirstlane(it.kind_inst);
        it.a = __builtin_amdgcn_readfirstlane(it.a);
        it.b = __builtin_amdgcn_readfirstlane(it.b);
        it.need = __builtin_amdgcn_readfirstlane(it.need);
        const int kind = it.kind_inst & 7, i0 = it.kind_inst >> 3;
        const bool multi = kind == kPLeaf || kind == kPExtract || kind == kPFinal;
        // inputs: wait for the instance's (instances') earlier phases (the
        // waiting lanes are in wave 0), then read through a clean L1
        if (kind == kPExtract || kind == kPFinal) {
            const int i = i0 + t;
            if (i <= it.need)
                persist_wait(&A.ctr[i], kind == kPExtract ? A.need_ex[i] : A.need_fin[i], &A.sched[kPSchedErr],
                             A.spin_limit);
        } else if (!multi && t == 0 && it.need > 0) {
            persist_wait(&A.ctr[i0], it.need, &A.sched[kPSchedErr], A.spin_limit);
        }
        if (kind != kPLeaf && t < 64) cache_invalidate();
        __syncthreads();
        switch (kind) {
        case kPLeaf:
            persist_leaf<B>(A.inst, A.blk_inst, A.S, A.blocks, A.NB, A.qend[2 * x], A.nq, A.q0, it.a);
            break;
        case kPBuildLds: {
            const LevelJob J = A.jobs[it.a];
            const int npair = (J.nr_parent + 1) >> 1;
            const FlatMap M{it.b, npair, J.nq * npair};
#define ICM_BUILD_F(FF) build_lds_body<FF, FlatMap, kLT, true>(J, lds, M, A.tw, A.twShift)
            ICM_F_CASES(ICM_BUILD_F)
#undef ICM_BUILD_F
            break;
        }
        case kPPropLds: {
            const LevelJob J = A.jobs[it.a];
            const int npair = (J.np + 1) >> 1;
            const FlatMap M{it.b, npair, J.nq * npair};
#define ICM_PROP_F(FF) prop_lds_body<FF, FlatMap, true>(J, lds, M, A.tw, A.twShift)
            ICM_F_CASES(ICM_PROP_F)
#undef ICM_PROP_F
            break;
        }
        case kPBuildSb: {
            const LevelJob J = A.jobs[it.a];
            build_sb_body<FlatMap, kLT, true>(J, reinterpret_cast<double *>(lds), FlatMap{it.b, J.nr_parent,
                                        J.nq * J.nr_parent}, sb_ppw(sb_stride_build(J.child.count)));
            break;
        }
        case kPPropSb: {
            const LevelJob J = A.jobs[it.a];
            prop_sb_body<FlatMap, kLT, true>(J, reinterpret_cast<double *>(lds), FlatMap{it.b, J.np, J.nq * J.np},
                                       sb_ppw(sb_stride_prop(J.child.count, J.gchild.count)));
            break;
        }
        case kPExtract:
            persist_extract<B>(A.inst, A.blk_inst, A.S, A.blocks, A.NB, A.qend[2 * x], A.nq, A.q0, A.q_per_chunk,
                               A.partial, A.Ntot, it.a, it.b);
            break;
        case kPFinal:
            persist_final(A, it.a, A.qend[2 * x + 1]);
            break;
        default:
            break;
        }
        if (kind != kPFinal) {   // nothing waits for finalize items
            pend = it.kind_inst;
            pend_i1 = it.need;
        }
    }
    // the last workgroup to leave checks that every queue was served and
    // clears the counters for the next launch
    const int t = threadIdx.x;
    if (t == 0) {
        const int e = __hip_atomic_fetch_add(&A.sched[kPSchedDone], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        s_last = (e == (int)gridDim.x - 1) ? 1 : 0;
    }
    __syncthreads();
    if (__builtin_amdgcn_readfirstlane(s_last)) {
        if (t < A.nqueues) {
            const int served = __hip_atomic_load(&A.sched[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (served < A.qoff[t + 1] - A.qoff[t])
                __hip_atomic_store(&A.sched[kPSchedErr], 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        for (int i = t; i < A.ninst; i += kLT) A.ctr[i] = 0;
        __syncthreads();
        if (t < A.nqueues) A.sched[t] = 0;
        if (t == 0) A.sched[kPSchedDone] = 0;
    }
}

// ---------------------------------------------------------------------------
// Fused adjoint levels, LDS-resident (F <= 4096), one parent per item.
//   build:  P_X = P_L P_R,  dP_X = dP_L P_R + P_L dP_R
//           FFT(P_L + iP_R), FFT(dP_L + i dP_R), inverse of P_X + i dP_X
//   propagate (A = forward g, B = adjoint of P):
//           A_L = A*P_R, A_R = A*P_L, B_L = B*P_R + A*dP_R, B_R = B*P_L + A*dP_L
//           FFT(A + iB), FFT(P_L + iP_R), FFT(dP_L + i dP_R),
//           inverses of A_L + iA_R and B_L + iB_R
// (* = the correlation of the propagate).  A missing right child has P = 1,
// dP = 0.  Coefficient fixes as the forward kernels: exact 0 / top terms in
// the build, the one wrapped B coefficient in the propagate.
// ---------------------------------------------------------------------------
// Branch-free, as load_child_pair.
__device__ __forceinline__ double2 load_view_pair(const CoefView &V, int qi, int p, int m, int nr_child,
                                                  bool missing_is_one) {
    const bool has_r = (2 * p + 1) < nr_child;
    const bool in = p >= 0 && m < V.count;
    const bool one = p >= 0 && !has_r && m == 0 && missing_is_one;
    const double *b = V.base + (long long)qi * V.plane + (long long)(2 * p) * V.stride + m;
    const double *px = in ? b : kZeroOne;
    const double *py = (in && has_r) ? b + V.stride : kZeroOne + (one ? 1 : 0);
    return make_double2(*as_global(px), *as_global(py));
}

// Item maps as the forward kernels, with one parent per item ("pair" = parent).
// LT threads per tile as the forward LDS kernels (512: radix 8, 256: radix 16).
template <int F, class Map, int LT = kLT>
__device__ __forceinline__ void build_lds_t_body(const LevelJob &J, double2 *lds, const Map &M, const double2 *tw,
                                                 int twShift) {
    using TL = typename LdsTL<F, LT>::type;
    constexpr int LP = kTile / LT;
    constexpr int HALF = F / 2;
    constexpr int LOGH = ilog2c(HALF);
    const int t = threadIdx.x;
    (void)twShift;
    // forward transforms take their first pass straight from global memory,
    // the inverse stores its last straight to P_X and dP_X
    fft_tile_io<F, false, LT, TL>(
        lds, tw,
        [&](int item, int m) {
            int qi, pr;
            M.at(item, qi, pr);
            return load_view_pair(J.child, qi, (pr < J.nr_parent) ? pr : -1, m, J.nr_child, true);
        },
        NoIO{});
    double2 XL[LP / 2], XR[LP / 2];
#pragma unroll
    for (int u = 0; u < LP / 2; ++u) {
        const int ps = t + u * LT;
        const int item = ps >> LOGH, jb = ps & (HALF - 1);
        const double2 za = lds[TL::at(item, jb)], zb = lds[TL::at(item, jb == 0 ? HALF : F - jb)];
        if (jb == 0) { XL[u] = make_double2(za.x, zb.x); XR[u] = make_double2(za.y, zb.y); }
        else { XL[u] = split_x(za, zb); XR[u] = split_y(za, zb); }
    }
    __syncthreads();
    fft_tile_io<F, false, LT, TL>(
        lds, tw,
        [&](int item, int m) {
            int qi, pr;
            M.at(item, qi, pr);
            return load_view_pair(J.tchild, qi, (pr < J.nr_parent) ? pr : -1, m, J.nr_child, false);
        },
        NoIO{});
#pragma unroll
    for (int u = 0; u < LP / 2; ++u) {
        const int ps = t + u * LT;
        const int item = ps >> LOGH, jb = ps & (HALF - 1);
        const int ia = TL::at(item, jb), ib = TL::at(item, jb == 0 ? HALF : F - jb);
        const double2 za = lds[ia], zb = lds[ib];
        if (jb == 0) {
            lds[ia] = make_double2(XL[u].x * XR[u].x, za.x * XR[u].x + XL[u].x * za.y);
            lds[ib] = make_double2(XL[u].y * XR[u].y, zb.x * XR[u].y + XL[u].y * zb.y);
        } else {
            const double2 DL = split_x(za, zb), DR = split_y(za, zb);
            const double2 Pv = c_mul(XL[u], XR[u]);
            const double2 Dv = c_add(c_mul(DL, XR[u]), c_mul(XL[u], DR));
            lds[ia] = pack_ai(Pv, Dv);
            lds[ib] = pack_ai(c_conj(Pv), c_conj(Dv));
        }
    }
    const double invF = 1.0 / (double)F;
    const int pc = J.parent.count, cc = J.child.count;
    fft_tile_io<F, true, LT, TL>(lds, tw, NoIO{}, [&](int item, int m, double2 v) {
        int qi, p;
        M.at(item, qi, p);
        if (p >= J.nr_parent) return;
        gdouble_rw *pp =
            as_global_rw(J.parent.base + (long long)qi * J.parent.plane + (long long)p * J.parent.stride);
        gdouble_rw *dp =
            as_global_rw(J.tparent.base + (long long)qi * J.tparent.plane + (long long)p * J.tparent.stride);
        const bool has_r = (2 * p + 1) < J.nr_child;
        auto exact = [&](int mm, double &pv, double &dv) {
            const double l = view_at(J.child, qi, 2 * p, mm), dl = view_at(J.tchild, qi, 2 * p, mm);
            const double r = has_r ? view_at(J.child, qi, 2 * p + 1, mm) : (mm == 0 ? 1.0 : 0.0);
            const double dr = has_r ? view_at(J.tchild, qi, 2 * p + 1, mm) : 0.0;
            pv = l * r;
            dv = dl * r + l * dr;
        };
        if (m < pc) {
            double pv, dv;
            if (m == 0) {
                exact(0, pv, dv);
            } else {
                pv = v.x * invF;
                dv = v.y * invF;
            }
            pp[m] = pv;
            dp[m] = dv;
        }
        if (m == 0 && pc > F) {
            double pv, dv;
            exact(cc - 1, pv, dv);
            pp[F] = pv;
            dp[F] = dv;
        }
    });
}

// Propagate order: (1) A + iB of the parent, whose spectra GA, GB stay in
// registers; (2) P_L + iP_R: the A outputs' spectra go to the tile and their
// inverse runs at once, while the B outputs keep their first terms
// B conj(P_R), B conj(P_L); (3) dP_L + i dP_R adds A conj(dP_R), A conj(dP_L)
// and the B inverse runs.  Three spectra per representative stay live (GA
// and the two B partial sums) instead of six, which lets the kernel run
// radix-16 passes at 256 threads without spills.  Each B coefficient is the
// same two products summed in the same order as computing both at once.
template <int F, class Map, int LT = kLT>
__device__ __forceinline__ void prop_lds_t_body(const LevelJob &J, double2 *lds, const Map &M, const double2 *tw,
                                                int twShift) {
    using TL = typename LdsTL<F, LT>::type;
    constexpr int LP = kTile / LT;
    constexpr int HALF = F / 2;
    constexpr int LOGH = ilog2c(HALF);
    const int t = threadIdx.x;
    (void)twShift;
    // (1) A + iB of the parent (every forward transform takes its first pass
    // straight from global memory, every inverse stores its last straight to
    // the outputs)
    fft_tile_io<F, false, LT, TL>(
        lds, tw,
        [&](int item, int m) {
            int qi, p;
            M.at(item, qi, p);
            double2 z = make_double2(0.0, 0.0);
            if (p < J.nr_parent) {
                if (m < J.gpar.count) z.x = view_at(J.gpar, qi, p, m);
                if (J.bpar.base && m < J.bpar.count) z.y = view_at(J.bpar, qi, p, m);
            }
            return z;
        },
        NoIO{});
    double2 GA[LP / 2], GB[LP / 2];
#pragma unroll
    for (int u = 0; u < LP / 2; ++u) {
        const int ps = t + u * LT;
        const int item = ps >> LOGH, jb = ps & (HALF - 1);
        const double2 za = lds[TL::at(item, jb)], zb = lds[TL::at(item, jb == 0 ? HALF : F - jb)];
        if (jb == 0) { GA[u] = make_double2(za.x, zb.x); GB[u] = make_double2(za.y, zb.y); }
        else { GA[u] = split_x(za, zb); GB[u] = split_y(za, zb); }
    }
    __syncthreads();
    // (2) P_L + iP_R: A outputs to the tile, first terms of the B outputs
    // (bins 0 and HALF, real: BL1 = (B conj(R), B conj(L)) at bin 0 and BR1
    // at bin HALF; other bins: BL1 = B conj(R), BR1 = B conj(L))
    fft_tile_io<F, false, LT, TL>(
        lds, tw,
        [&](int item, int m) {
            int qi, pr;
            M.at(item, qi, pr);
            return load_view_pair(J.child, qi, (pr < J.nr_parent) ? pr : -1, m, J.nr_child, true);
        },
        NoIO{});
    double2 BL1[LP / 2], BR1[LP / 2];
#pragma unroll
    for (int u = 0; u < LP / 2; ++u) {
        const int ps = t + u * LT;
        const int item = ps >> LOGH, jb = ps & (HALF - 1);
        const int ia = TL::at(item, jb), ib = TL::at(item, jb == 0 ? HALF : F - jb);
        const double2 za = lds[ia], zb = lds[ib];
        if (jb == 0) {
            // bins 0 and HALF carry real values (x: bin 0, y: bin HALF)
            const double2 XL = make_double2(za.x, zb.x), XR = make_double2(za.y, zb.y);
            lds[ia] = make_double2(GA[u].x * XR.x, GA[u].x * XL.x);
            lds[ib] = make_double2(GA[u].y * XR.y, GA[u].y * XL.y);
            BL1[u] = make_double2(GB[u].x * XR.x, GB[u].x * XL.x);
            BR1[u] = make_double2(GB[u].y * XR.y, GB[u].y * XL.y);
        } else {
            const double2 XL = split_x(za, zb), XR = split_y(za, zb);
            const double2 AL = c_mulc(GA[u], XR), AR = c_mulc(GA[u], XL);
            lds[ia] = pack_ai(AL, AR);
            lds[ib] = pack_ai(c_conj(AL), c_conj(AR));
            BL1[u] = c_mulc(GB[u], XR);
            BR1[u] = c_mulc(GB[u], XL);
        }
    }
    const double invF = 1.0 / (double)F;
    fft_tile_io<F, true, LT, TL>(lds, tw, NoIO{}, [&](int item, int m, double2 v) {
        int qi, p;
        M.at(item, qi, p);
        if (p < J.nr_parent && m < J.gchild.count) {
            gdouble_rw *g = as_global_rw(J.gchild.base + (long long)qi * J.gchild.plane +
                                         (long long)(2 * p) * J.gchild.stride);
            g[m] = v.x * invF;
            if (2 * p + 1 < J.nr_child) g[J.gchild.stride + m] = v.y * invF;
        }
    });
    // (3) dP_L + i dP_R: second terms of the B outputs, their inverse
    fft_tile_io<F, false, LT, TL>(
        lds, tw,
        [&](int item, int m) {
            int qi, pr;
            M.at(item, qi, pr);
            return load_view_pair(J.tchild, qi, (pr < J.nr_parent) ? pr : -1, m, J.nr_child, false);
        },
        NoIO{});
#pragma unroll
    for (int u = 0; u < LP / 2; ++u) {
        const int ps = t + u * LT;
        const int item = ps >> LOGH, jb = ps & (HALF - 1);
        const int ia = TL::at(item, jb), ib = TL::at(item, jb == 0 ? HALF : F - jb);
        const double2 za = lds[ia], zb = lds[ib];
        if (jb == 0) {
            const double dl0 = za.x, dr0 = za.y, dlh = zb.x, drh = zb.y;
            lds[ia] = make_double2(BL1[u].x + GA[u].x * dr0, BL1[u].y + GA[u].x * dl0);
            lds[ib] = make_double2(BR1[u].x + GA[u].y * drh, BR1[u].y + GA[u].y * dlh);
        } else {
            const double2 DL = split_x(za, zb), DR = split_y(za, zb);
            const double2 BL = c_add(BL1[u], c_mulc(GA[u], DR));
            const double2 BR = c_add(BR1[u], c_mulc(GA[u], DL));
            lds[ia] = pack_ai(BL, BR);
            lds[ib] = pack_ai(c_conj(BL), c_conj(BR));
        }
    }
    const int cc = J.child.count;
    fft_tile_io<F, true, LT, TL>(lds, tw, NoIO{}, [&](int item, int m, double2 v) {
        int qi, p;
        M.at(item, qi, p);
        if (p < J.nr_parent && m < J.bchild.count) {
            double bl = v.x * invF, br = v.y * invF;
            if (J.gfix && m == cc - 1) {
                // linear correlation reads Y[F], the cyclic one read Y[0]
                const bool has_r = (2 * p + 1) < J.nr_child;
                const double aF = (F < J.gpar.count) ? view_at(J.gpar, qi, p, F) : 0.0;
                const double a0 = view_at(J.gpar, qi, p, 0);
                const double bF = (J.bpar.base && F < J.bpar.count) ? view_at(J.bpar, qi, p, F) : 0.0;
                const double b0 = J.bpar.base ? view_at(J.bpar, qi, p, 0) : 0.0;
                const double pl = view_at(J.child, qi, 2 * p, cc - 1), dl = view_at(J.tchild, qi, 2 * p, cc - 1);
                const double pr = has_r ? view_at(J.child, qi, 2 * p + 1, cc - 1) : (cc == 1 ? 1.0 : 0.0);
                const double dr = has_r ? view_at(J.tchild, qi, 2 * p + 1, cc - 1) : 0.0;
                bl += pr * (bF - b0) + dr * (aF - a0);
                br += pl * (bF - b0) + dl * (aF - a0);
            }
            gdouble_rw *bb = as_global_rw(J.bchild.base + (long long)qi * J.bchild.plane +
                                          (long long)(2 * p) * J.bchild.stride);
            bb[m] = bl;
            if (2 * p + 1 < J.nr_child) bb[J.bchild.stride + m] = br;
        }
    });
}


// Threads per tile as the forward LDS kernels; two workgroups per CU.
template <int F> constexpr int lds_t_threads() { return lds_build_threads<F>(); }
template <int F>
__global__ __launch_bounds__(lds_t_threads<F>(), lds_t_threads<F>() == 256 ? 2 : 4) void k_build_lds_t(
    LevelJob J, const double2 *tw, int twShift) {
    __shared__ double2 lds[kTile];
    build_lds_t_body<F, PlainMap, lds_t_threads<F>()>(
        J, lds, PlainMap{(int)blockIdx.x * (kTile / F), (int)blockIdx.y}, tw, twShift);
}

template <int F> constexpr int lds_prop_t_threads() { return lds_build_threads<F>(); }
template <int F>
__global__ __launch_bounds__(lds_prop_t_threads<F>(), lds_prop_t_threads<F>() == 256 ? 2 : 4) void k_prop_lds_t(
    LevelJob J, const double2 *tw, int twShift) {
    __shared__ double2 lds[kTile];
    prop_lds_t_body<F, PlainMap, lds_prop_t_threads<F>()>(
        J, lds, PlainMap{(int)blockIdx.x * (kTile / F), (int)blockIdx.y}, tw, twShift);
}

// Ragged batches, adjoint levels: work item = (job, first parent item).
__global__ __launch_bounds__(512) void k_build_lds_t_batch(const LevelJob *__restrict__ jobs,
                                                           const int2 *__restrict__ items, const double2 *tw,
                                                           int twShift) {
    __shared__ double2 lds[kTileMaxSize];
    const int2 it = items[blockIdx.x];
    const LevelJob J = jobs[it.x];
    const FlatMap M{it.y, J.nr_parent, J.nq * J.nr_parent};
#define ICM_BUILD_T_F(FF) build_lds_t_body<FF>(J, lds, M, tw, twShift)
    ICM_F_CASES(ICM_BUILD_T_F)
#undef ICM_BUILD_T_F
}

__global__ __launch_bounds__(512) void k_prop_lds_t_batch(const LevelJob *__restrict__ jobs,
                                                          const int2 *__restrict__ items, const double2 *tw,
                                                          int twShift) {
    __shared__ double2 lds[kTileMaxSize];
    const int2 it = items[blockIdx.x];
    const LevelJob J = jobs[it.x];
    const FlatMap M{it.y, J.nr_parent, J.nq * J.nr_parent};
#define ICM_PROP_T_F(FF) prop_lds_t_body<FF>(J, lds, M, tw, twShift)
    ICM_F_CASES(ICM_PROP_T_F)
#undef ICM_PROP_T_F
}

// ---------------------------------------------------------------------------
// Large levels (F >= 8192): four-step transform with F = F1 x F2, F2 = 4096.
//   forward: column FFTs (size F1, stride F2) + twiddle w_F^(n2 k1) -> [k1][n2]
//            row FFTs (size F2) -> spectrum in [k1][k2] order (K = k1 + F1 k2)
//   inverse: row IFFTs of [k1][k2] -> [k1][n2]
//            twiddle w_F^(-n2 k1) (on the column kernel's load) + column
//            IFFTs -> natural order.
// Spectra stay in [k1][k2] order; the pointwise kernels find -K there.
// ---------------------------------------------------------------------------
constexpr int kF2 = kTile;

// Two coefficient rows (x = Re, y = Im) of one transform, resolved once per
// transform: the loaders' parent lookups and row addresses stay out of the
// per-point code.  Point m: (x[m], y[m]) below the counts, 0 past them; with
// no y row, y[0] = `one` (a missing right child as the identity, or 0).
struct RowPair {
    const double *x, *y;
    int cx, cy;
    double one;   // 1: the y row is the identity polynomial (only without a y row)
    // branch-free, as load_child_pair
    __device__ __forceinline__ double2 at(int m) const {
        const double *px = m < cx ? x + m : kZeroOne;
        const double *py = m < cy ? y + m : kZeroOne + ((m == 0 && one != 0.0) ? 1 : 0);
        return make_double2(*px, *py);
    }
};
// Child pair of slot a: left child (x) and right child (y) of its parent, as
// load_child_pair gives them.
struct ChildRows {
    LevelJob J;
    __device__ __forceinline__ RowPair rows(int a, int qi) const {
        RowPair r{nullptr, nullptr, 0, 0, 0.0};
        const int p = slot_parent(J, a);
        if (p < 0 || p >= J.nr_parent) return r;
        const bool has_r = (2 * p + 1) < J.nr_child;
        r.x = J.child.base + (long long)qi * J.child.plane + (long long)(2 * p) * J.child.stride;
        r.cx = J.child.count;
        if (has_r) {
            const CoefView &R = J.child_r.base ? J.child_r : J.child;
            r.y = R.base + (long long)qi * R.plane + (long long)(2 * p + 1) * R.stride;
            r.cy = J.child.count;
        } else if (!J.r_missing_zero) {
            r.one = 1.0;
        }
        return r;
    }
};
// g pair `a`: slots 2a (x) and 2a + 1 (y) of the parents' g.
struct GRows {
    LevelJob J;
    __device__ __forceinline__ RowPair rows(int a, int qi) const {
        RowPair r{nullptr, nullptr, 0, 0, 0.0};
        const double *gb = J.gpar.base + (long long)qi * J.gpar.plane;
        const int pa = slot_parent(J, 2 * a), pb = slot_parent(J, 2 * a + 1);
        if (pa >= 0) {
            r.x = gb + (long long)pa * J.gpar.stride;
            r.cx = J.gpar.count;
        }
        if (pb >= 0) {
            r.y = gb + (long long)pb * J.gpar.stride;
            r.cy = J.gpar.count;
        }
        return r;
    }
};
// Storers of the inverse column stage.  rows(a, qi) resolves array a's
// output rows once (uniform over a workgroup); store(o, qi, n, v) writes
// coefficient n of them from v = (x, y) before the 1/F scale.
struct OutRows {
    double *x, *y;   // null: no such node
    int p;           // first node of the array
};
struct StParentPair {     // array a = pair of parents (build)
    LevelJob J;
    __device__ OutRows rows(int a, int qi) const {
        double *pb = J.parent.base + (long long)qi * J.parent.plane;
        const int p = 2 * a;
        return OutRows{p < J.nr_parent ? pb + (long long)p * J.parent.stride : nullptr,
                       p + 1 < J.nr_parent ? pb + (long long)(p + 1) * J.parent.stride : nullptr, p};
    }
    __device__ void store(const OutRows &o, int qi, int n, double2 v) const {
        const double invF = 1.0 / (double)J.F;
        const int pc = J.parent.count;
        if (n < pc) {
            if (o.x) put(&o.x[n], (n == 0) ? child_prod_at(J, qi, o.p, 0) : v.x * invF, J.acc_out);
            if (o.y) put(&o.y[n], (n == 0) ? child_prod_at(J, qi, o.p + 1, 0) : v.y * invF, J.acc_out);
        }
        if (n == 0 && pc > J.F) {
            const int top = J.child.count - 1;
            if (o.x) put(&o.x[J.F], child_prod_at(J, qi, o.p, top), J.acc_out);
            if (o.y) put(&o.y[J.F], child_prod_at(J, qi, o.p + 1, top), J.acc_out);
        }
    }
};
struct StParentTangent {  // array a = pair of parents (adjoint tangent build): dP_X into J.parent
    LevelJob J;              // child = P of the children, tchild = their dP
    __device__ double exact(int qi, int p, int m) const {   // dP_L[m] P_R[m] + P_L[m] dP_R[m]
        const double pl = view_at(J.child, qi, 2 * p, m), dl = view_at(J.tchild, qi, 2 * p, m);
        const bool has_r = (2 * p + 1) < J.nr_child;
        const double pr = has_r ? view_at(J.child, qi, 2 * p + 1, m) : (m == 0 ? 1.0 : 0.0);
        const double dr = has_r ? view_at(J.tchild, qi, 2 * p + 1, m) : 0.0;
        return dl * pr + pl * dr;
    }
    __device__ OutRows rows(int a, int qi) const {
        double *pb = J.parent.base + (long long)qi * J.parent.plane;
        const int p = 2 * a;
        return OutRows{p < J.nr_parent ? pb + (long long)p * J.parent.stride : nullptr,
                       p + 1 < J.nr_parent ? pb + (long long)(p + 1) * J.parent.stride : nullptr, p};
    }
    __device__ void store(const OutRows &o, int qi, int n, double2 v) const {
        const double invF = 1.0 / (double)J.F;
        const int pc = J.parent.count;
        if (n < pc) {
            if (o.x) put(&o.x[n], (n == 0) ? exact(qi, o.p, 0) : v.x * invF, J.acc_out);
            if (o.y) put(&o.y[n], (n == 0) ? exact(qi, o.p + 1, 0) : v.y * invF, J.acc_out);
        }
        if (n == 0 && pc > J.F) {
            const int top = J.child.count - 1;
            if (o.x) put(&o.x[J.F], exact(qi, o.p, top), J.acc_out);
            if (o.y) put(&o.y[J.F], exact(qi, o.p + 1, top), J.acc_out);
        }
    }
};
struct StGChildren {      // array a = slot (propagate): g of the slot's parent's two children
    LevelJob J;
    __device__ OutRows rows(int a, int qi) const {
        const int p = slot_parent(J, a);
        if (p < 0) return OutRows{nullptr, nullptr, -1};
        double *l = J.gchild.base + (long long)qi * J.gchild.plane + (long long)(2 * p) * J.gchild.stride;
        return OutRows{l, (2 * p + 1 < J.nr_child) ? l + J.gchild.stride : nullptr, p};
    }
    __device__ void store(const OutRows &o, int qi, int n, double2 v) const {
        if (!o.x || n >= J.gchild.count) return;
        const double invF = 1.0 / (double)J.F;
        double vl = v.x * invF, vr = v.y * invF;
        if (J.gfix && n == J.child.count - 1) {
            vl += prop_wrap_term(J, qi, o.p, 0);
            vr += prop_wrap_term(J, qi, o.p, 1);
        }
        put(&o.x[n], vl, J.acc_out);
        if (o.y) put(&o.y[n], vr, J.acc_out);
    }
};
// Sum of two propagates into the same children (J1 with its parents' g and
// children, J2 likewise; J1's outputs, counts and flags): the one wrapped
// coefficient of each correlation is fixed exactly.
struct StGChildren2 {
    LevelJob J1, J2;
    __device__ OutRows rows(int a, int qi) const { return StGChildren{J1}.rows(a, qi); }
    __device__ void store(const OutRows &o, int qi, int n, double2 v) const {
        if (!o.x || n >= J1.gchild.count) return;
        const double invF = 1.0 / (double)J1.F;
        double vl = v.x * invF, vr = v.y * invF;
        if (J1.gfix && n == J1.child.count - 1) {
            vl += prop_wrap_term(J1, qi, o.p, 0) + prop_wrap_term(J2, qi, o.p, 0);
            vr += prop_wrap_term(J1, qi, o.p, 1) + prop_wrap_term(J2, qi, o.p, 1);
        }
        put(&o.x[n], vl, J1.acc_out);
        if (o.y) put(&o.y[n], vr, J1.acc_out);
    }
};

// Column kernels: 512 threads (8 points each, radix-8 passes), so two 64 KB
// workgroups keep 16 waves per CU in flight for the HBM-bound passes.
constexpr int kCT = 512;
constexpr int kCP = kTile / kCT;

template <int F1, class Rows>
__global__ __launch_bounds__(kCT) void k_col_fwd(Rows src, double2 *__restrict__ out, int nArr, int logF,
                                                 const double2 *tw, int twShift) {
    using TL = ColLayout<F1>;
    __shared__ double2 lds[TL::SIZE];
    constexpr int TC = kTile / F1;
    constexpr int LOGTC = ilog2c(TC);
    const int tile = blockIdx.x, a = blockIdx.y, qi = blockIdx.z;
    const int t = threadIdx.x;
    const int F = 1 << logF;
    const RowPair rp = src.rows(a, qi);   // uniform over the workgroup
#pragma unroll
    for (int u = 0; u < kCP; ++u) {
        const int e = t + u * kCT;
        const int r = e >> LOGTC, c = e & (TC - 1);
        const int n = r * kF2 + tile * TC + c;
        lds[TL::at(c, r)] = rp.at(n);
    }
    fft_tile<F1, false, kCT, TL>(lds, tw, twShift);
    double2 *o = out + ((long long)qi * nArr + a) * F;
    if constexpr (TC <= kCT) {
        // A thread's outputs share column n2 at rows r0 + u RS: its twiddles
        // are w_F^(n2 r0) times powers of w_F^(n2 RS), two table lookups
        // instead of eight (the lookups gather from scattered table lines;
        // the product chain adds at most one rounding per step).
        constexpr int RS = kCT / TC;
        const int c = t & (TC - 1), r0 = t >> LOGTC;
        const int n2 = tile * TC + c;
        double2 w = tw_big(tw, twShift, (n2 * r0) & (F - 1), logF);
        const double2 ws = tw_big(tw, twShift, (n2 * RS) & (F - 1), logF);
#pragma unroll
        for (int u = 0; u < kCP; ++u) {
            const int r = r0 + u * RS;
            o[r * kF2 + n2] = c_mul(lds[TL::at(c, r)], w);
            w = c_mul(w, ws);
        }
    } else {
#pragma unroll
        for (int u = 0; u < kCP; ++u) {
            const int e = t + u * kCT;
            const int r = e >> LOGTC, c = e & (TC - 1);
            const int n2 = tile * TC + c;
            const double2 w = tw_big(tw, twShift, (n2 * r) & (F - 1), logF);
            o[r * kF2 + n2] = c_mul(lds[TL::at(c, r)], w);
        }
    }
}

template <int F1, class Storer>
__global__ __launch_bounds__(kCT) void k_col_inv(const double2 *__restrict__ in, Storer st, int nArr, int logF,
                                                 const double2 *tw, int twShift) {
    using TL = ColLayout<F1>;
    __shared__ double2 lds[TL::SIZE];
    constexpr int TC = kTile / F1;
    constexpr int LOGTC = ilog2c(TC);
    const int tile = blockIdx.x, a = blockIdx.y, qi = blockIdx.z;
    const int t = threadIdx.x;
    const int F = 1 << logF;
    const double2 *src = in + ((long long)qi * nArr + a) * F;
    // the inverse four-step twiddle w_F^(-n2 r) of the row stage (powers of
    // one step per thread where its inputs share a column, as in k_col_fwd)
    if constexpr (TC <= kCT) {
        constexpr int RS = kCT / TC;
        const int c = t & (TC - 1), r0 = t >> LOGTC;
        const int n2 = tile * TC + c;
        double2 w = tw_big2(tw, twShift, (n2 * r0) & (F - 1), logF);
        const double2 ws = tw_big2(tw, twShift, (n2 * RS) & (F - 1), logF);
#pragma unroll
        for (int u = 0; u < kCP; ++u) {
            const int r = r0 + u * RS;
            lds[TL::at(c, r)] = c_mulc(src[r * kF2 + n2], w);
            w = c_mul(w, ws);
        }
    } else {
#pragma unroll
        for (int u = 0; u < kCP; ++u) {
            const int e = t + u * kCT;
            const int r = e >> LOGTC, c = e & (TC - 1);
            const int n2 = tile * TC + c;
            lds[TL::at(c, r)] = c_mulc(src[r * kF2 + n2], tw_big2(tw, twShift, (n2 * r) & (F - 1), logF));
        }
    }
    fft_tile<F1, true, kCT, TL>(lds, tw, twShift);
    const OutRows o = st.rows(a, qi);   // uniform over the workgroup
#pragma unroll
    for (int u = 0; u < kCP; ++u) {
        const int e = t + u * kCT;
        const int r = e >> LOGTC, c = e & (TC - 1);
        st.store(o, qi, r * kF2 + tile * TC + c, lds[TL::at(c, r)]);
    }
}

// Propagate column inverse of a four-step level l fused with the g-pair
// column forward of level l - 1 (also four-step, F' = F / 2).  The inverse
// of slot a holds (g_left, g_right) of parent a's children as (Re, Im), which
// is exactly level l-1's g pair a; its coefficients below F / 2 = F' are
// rows r < F1 / 2 of every column, so level l-1's column n2 is the first
// half of level l's column n2.  A workgroup inverts two column tiles of level
// l (2 TC columns, one tile at a time, keeping the first-half rows in
// registers), then runs level l-1's forward columns on one tile of
// (2 TC columns x F1/2 rows) and writes them as k_col_fwd<F1/2, GRows> does
// (same values: the scaled coefficients, zero past the count and for a
// missing right child, then the same passes and twiddles).  Level l-1's g
// coefficients are never stored.
// keep_g0: coefficient 0 of every child's g is stored as well (the adjoint's
// B propagate of the level below reads it for its wrapped term).
// B2: the adjoint's one-pass B propagate instead (StGChildren2's values: J's
// and J2's wrapped terms added at coefficient c - 1), storing each child's
// coefficients 0 and F / 2, which the level below's wrapped terms read.
template <int F1, bool B2 = false>
__global__ __launch_bounds__(kCT) void k_col_inv_fwd(const double2 *__restrict__ in, LevelJob J, LevelJob J2,
                                                     double2 *__restrict__ out, int nArr, int logF,
                                                     const double2 *tw, int twShift, int keep_g0) {
    static_assert(F1 >= 8, "the next level must be a four-step level with column kernels");
    using TL = ColLayout<F1>;
    constexpr int H = F1 / 2;
    using TLH = ColLayout<H>;
    __shared__ double2 lds[kTile];
    constexpr int TC = kTile / F1;
    constexpr int LOGTC = ilog2c(TC);
    constexpr int KEEP = kTile / 2 / kCT;     // first-half points per thread and tile
    const int a = blockIdx.y, qi = blockIdx.z, t = threadIdx.x;
    const int F = 1 << logF;
    const double2 *src = in + ((long long)qi * nArr + a) * F;
    const int p = slot_parent(J, a);
    const bool has_l = p >= 0, has_r = p >= 0 && (2 * p + 1) < J.nr_child;
    const int gcount = J.gchild.count;
    const double invF = 1.0 / (double)F;
    double2 held[2][KEEP];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int tile = 2 * blockIdx.x + h;
        if constexpr (TC <= kCT) {   // inverse four-step twiddle on the load, as k_col_inv
            constexpr int RS = kCT / TC;
            const int c = t & (TC - 1), r0 = t >> LOGTC;
            const int n2 = tile * TC + c;
            double2 w = tw_big2(tw, twShift, (n2 * r0) & (F - 1), logF);
            const double2 ws = tw_big2(tw, twShift, (n2 * RS) & (F - 1), logF);
#pragma unroll
            for (int u = 0; u < kCP; ++u) {
                const int r = r0 + u * RS;
                lds[TL::at(c, r)] = c_mulc(src[r * kF2 + n2], w);
                w = c_mul(w, ws);
            }
        } else {
#pragma unroll
            for (int u = 0; u < kCP; ++u) {
                const int e = t + u * kCT;
                const int r = e >> LOGTC, c = e & (TC - 1);
                const int n2 = tile * TC + c;
                lds[TL::at(c, r)] = c_mulc(src[r * kF2 + n2], tw_big2(tw, twShift, (n2 * r) & (F - 1), logF));
            }
        }
        fft_tile<F1, true, kCT, TL>(lds, tw, twShift);
        // coefficient n of both children from inverse value v (scaled; B2:
        // the wrapped terms at c - 1)
        auto value = [&](int n, double2 v) {
            double vx = v.x * invF, vy = v.y * invF;
            if constexpr (B2) {
                if (has_l && J.gfix && n == J.child.count - 1) {
                    vx += prop_wrap_term(J, qi, p, 0) + prop_wrap_term(J2, qi, p, 0);
                    vy += prop_wrap_term(J, qi, p, 1) + prop_wrap_term(J2, qi, p, 1);
                }
            }
            return make_double2((has_l && n < gcount) ? vx : 0.0, (has_r && n < gcount) ? vy : 0.0);
        };
        if constexpr (B2) {
            if (t < 2 && has_l) {
                const int n = t == 0 ? 0 : F / 2;
                if (n < gcount && ((n & (kF2 - 1)) >> LOGTC) == tile) {
                    const int r = n >> 12, c = (n & (kF2 - 1)) - tile * TC;
                    const double2 z = value(n, lds[TL::at(c, r)]);
                    double *gl = J.gchild.base + (long long)qi * J.gchild.plane + (long long)(2 * p) * J.gchild.stride;
                    gl[n] = z.x;
                    if (has_r) gl[J.gchild.stride + n] = z.y;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < KEEP; ++u) {
            const int e = t + u * kCT;            // e < kTile / 2: rows r < F1 / 2
            const int r = e >> LOGTC, c = e & (TC - 1);
            const int n = r * kF2 + tile * TC + c;
            held[h][u] = value(n, lds[TL::at(c, r)]);
            if (!B2 && keep_g0 && n == 0 && has_l) {
                double *gl = J.gchild.base + (long long)qi * J.gchild.plane + (long long)(2 * p) * J.gchild.stride;
                gl[0] = held[h][u].x;
                if (has_r) gl[J.gchild.stride] = held[h][u].y;
            }
        }
        __syncthreads();                          // reads done before the tile is reused
    }
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int u = 0; u < KEEP; ++u) {
            const int e = t + u * kCT;
            const int r = e >> LOGTC, c = e & (TC - 1);
            lds[TLH::at(h * TC + c, r)] = held[h][u];
        }
    fft_tile<H, false, kCT, TLH>(lds, tw, twShift);
    // level l-1's four-step twiddle w_F'^(n2 k1) and store, as k_col_fwd<H>
    constexpr int TC2 = 2 * TC;
    constexpr int LOGTC2 = ilog2c(TC2);
    const int Fh = F >> 1, logFh = logF - 1;
    double2 *o = out + ((long long)qi * nArr + a) * Fh;
    if constexpr (TC2 <= kCT) {
        constexpr int RS = kCT / TC2;
        const int c = t & (TC2 - 1), r0 = t >> LOGTC2;
        const int n2 = blockIdx.x * TC2 + c;
        double2 w = tw_big(tw, twShift, (n2 * r0) & (Fh - 1), logFh);
        const double2 ws = tw_big(tw, twShift, (n2 * RS) & (Fh - 1), logFh);
#pragma unroll
        for (int u = 0; u < kCP; ++u) {
            const int r = r0 + u * RS;
            o[r * kF2 + n2] = c_mul(lds[TLH::at(c, r)], w);
            w = c_mul(w, ws);
        }
    } else {
#pragma unroll
        for (int u = 0; u < kCP; ++u) {
            const int e = t + u * kCT;
            const int r = e >> LOGTC2, c = e & (TC2 - 1);
            const int n2 = blockIdx.x * TC2 + c;
            const double2 w = tw_big(tw, twShift, (n2 * r) & (Fh - 1), logFh);
            o[r * kF2 + n2] = c_mul(lds[TLH::at(c, r)], w);
        }
    }
}

// Build column inverse of a four-step level l fused with the child-pair
// column forward of level l + 1 (F' = 2 F).  Array a of the inverse holds
// parents 2a and 2a + 1 of level l as (Re, Im): the child pair of level
// l+1's parent a.  Level l+1's column n2 holds coefficients n2 + 4096 r',
// r' < 2 F1: rows r' < F1 are level l's inverse column n2 (scaled, with the
// exact coefficient 0 and zero past the count), row F1 of column 0 is the
// exact top coefficient (index F) when the parents keep F + 1 terms, and the
// rest is zero.  A workgroup inverts one column tile (TC columns), keeps it
// in registers and runs level l+1's forward columns on its two halves (TC/2
// columns x 2 F1 rows each), written as k_col_fwd<2 F1, ChildRows> writes
// them (same values, passes and twiddles).  Of level l's coefficients only
// the two end terms of every parent are stored (the level above computes its
// own end terms exactly from them); the rest are never stored: fused only
// where every later reader takes the spectra.
// TAN: the adjoint's tangent build instead (dP_X into J.parent, J.child the
// children's P, J.tchild their dP, as StParentTangent stores it), handing the
// level above its dP child-pair transforms (a missing right child 0).
template <int F1, bool TAN = false>
__global__ __launch_bounds__(kCT) void k_col_inv_fwd_build(const double2 *__restrict__ in, LevelJob J,
                                                           double2 *__restrict__ out, int nArr, int logF,
                                                           const double2 *tw, int twShift) {
    static_assert(F1 >= 4 && F1 <= 1024, "both levels are four-step levels with column kernels");
    using TL = ColLayout<F1>;
    constexpr int F1N = 2 * F1;
    using TLN = ColLayout<F1N>;
    __shared__ double2 lds[kTile];
    constexpr int TC = kTile / F1;
    constexpr int LOGTC = ilog2c(TC);
    constexpr int TCN = TC / 2;
    constexpr int LOGTCN = ilog2c(TCN);
    const int tile = blockIdx.x, a = blockIdx.y, qi = blockIdx.z, t = threadIdx.x;
    const int F = 1 << logF;
    const double2 *src = in + ((long long)qi * nArr + a) * F;
    if constexpr (TC <= kCT) {   // inverse four-step twiddle on the load, as k_col_inv
        constexpr int RS = kCT / TC;
        const int c = t & (TC - 1), r0 = t >> LOGTC;
        const int n2 = tile * TC + c;
        double2 w = tw_big2(tw, twShift, (n2 * r0) & (F - 1), logF);
        const double2 ws = tw_big2(tw, twShift, (n2 * RS) & (F - 1), logF);
#pragma unroll
        for (int u = 0; u < kCP; ++u) {
            const int r = r0 + u * RS;
            lds[TL::at(c, r)] = c_mulc(src[r * kF2 + n2], w);
            w = c_mul(w, ws);
        }
    } else {
#pragma unroll
        for (int u = 0; u < kCP; ++u) {
            const int e = t + u * kCT;
            const int r = e >> LOGTC, c = e & (TC - 1);
            const int n2 = tile * TC + c;
            lds[TL::at(c, r)] = c_mulc(src[r * kF2 + n2], tw_big2(tw, twShift, (n2 * r) & (F - 1), logF));
        }
    }
    fft_tile<F1, true, kCT, TL>(lds, tw, twShift);
    // level l's parent coefficients as StParentPair would store them
    const int p = 2 * a;
    const bool hx = p < J.nr_parent, hy = p + 1 < J.nr_parent;
    const int pc = J.parent.count;
    const double invF = 1.0 / (double)F;
    double *pb = J.parent.base + (long long)qi * J.parent.plane;
    double2 v[kCP];
#pragma unroll
    for (int u = 0; u < kCP; ++u) {
        const int e = t + u * kCT;
        const int r = e >> LOGTC, c = e & (TC - 1);
        const int n = r * kF2 + tile * TC + c;
        const double2 z = lds[TL::at(c, r)];
        double x = 0.0, y = 0.0;
        if (n == 0) {
            if constexpr (TAN) {
                if (hx) x = StParentTangent{J}.exact(qi, p, 0);
                if (hy) y = StParentTangent{J}.exact(qi, p + 1, 0);   // a missing right child's dP is 0
            } else {
                if (hx) x = child_prod_at(J, qi, p, 0);
                y = hy ? child_prod_at(J, qi, p + 1, 0) : 1.0;   // a missing right child is the identity
            }
        } else if (n < pc) {
            if (hx) x = z.x * invF;
            if (hy) y = z.y * invF;
        }
        v[u] = make_double2(x, y);
        // the level above's exact end terms read coefficients 0 and pc - 1
        if (n == 0 || n == pc - 1) {
            if (hx) pb[(long long)p * J.parent.stride + n] = x;
            if (hy) pb[(long long)(p + 1) * J.parent.stride + n] = y;
        }
    }
    double2 top = make_double2(0.0, 0.0);       // coefficient F, at row F1 of column 0
    if (tile == 0 && pc > F) {
        const int tc = J.child.count - 1;
        if constexpr (TAN) {
            if (hx) top.x = StParentTangent{J}.exact(qi, p, tc);
            if (hy) top.y = StParentTangent{J}.exact(qi, p + 1, tc);
        } else {
            if (hx) top.x = child_prod_at(J, qi, p, tc);
            if (hy) top.y = child_prod_at(J, qi, p + 1, tc);
        }
        if (t == 0) {
            if (hx) pb[(long long)p * J.parent.stride + F] = top.x;
            if (hy) pb[(long long)(p + 1) * J.parent.stride + F] = top.y;
        }
    }
    const int FN = F << 1, logFN = logF + 1;
    double2 *o = out + ((long long)qi * nArr + a) * FN;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        __syncthreads();                          // the previous contents are read
#pragma unroll
        for (int u = 0; u < kCP; ++u) {
            const int e = t + u * kCT;
            const int r = e >> LOGTC, c = e & (TC - 1);
            if ((c >> LOGTCN) == h) lds[TLN::at(c - h * TCN, r)] = v[u];
        }
#pragma unroll
        for (int u = 0; u < kCP / 2; ++u) {       // rows F1 .. 2 F1 - 1
            const int e = t + u * kCT;
            const int r = F1 + (e >> LOGTCN), c = e & (TCN - 1);
            lds[TLN::at(c, r)] = (h == 0 && c == 0 && r == F1) ? top : make_double2(0.0, 0.0);
        }
        fft_tile<F1N, false, kCT, TLN>(lds, tw, twShift);
        const int tn = 2 * tile + h;              // level l+1's column tile
        if constexpr (TCN <= kCT) {
            constexpr int RS = kCT / TCN;
            const int c = t & (TCN - 1), r0 = t >> LOGTCN;
            const int n2 = tn * TCN + c;
            double2 w = tw_big(tw, twShift, (n2 * r0) & (FN - 1), logFN);
            const double2 ws = tw_big(tw, twShift, (n2 * RS) & (FN - 1), logFN);
#pragma unroll
            for (int u = 0; u < kCP; ++u) {
                const int r = r0 + u * RS;
                o[r * kF2 + n2] = c_mul(lds[TLN::at(c, r)], w);
                w = c_mul(w, ws);
            }
        } else {
#pragma unroll
            for (int u = 0; u < kCP; ++u) {
                const int e = t + u * kCT;
                const int r = e >> LOGTCN, c = e & (TCN - 1);
                const int n2 = tn * TCN + c;
                const double2 w = tw_big(tw, twShift, (n2 * r) & (FN - 1), logFN);
                o[r * kF2 + n2] = c_mul(lds[TLN::at(c, r)], w);
            }
        }
    }
}

// Fused spectrum stage of a large level.  Every bin -K of row k1 lies in row
// F1 - k1 (position F2-1-k2; rows 0 and F1/2 are their own mirrors, at
// (F2-k2) mod F2 and F2-1-k2), so the forward row FFTs, the real-pair split,
// the pointwise product and the inverse row FFTs of a row and its mirror row
// happen without leaving the workgroup.
//
// Workgroup w of a level owns two rows, one 4096-point LDS tile each (128 KiB):
// w >= 1 the pair {w, F1 - w}; w = 0 the two self-mirrored rows {0, F1/2}.
// Both rows go through every pass together (one barrier per pass): threads
// 0..255 run tile 0 and 256..511 tile 1, radix-16 butterflies of 16 points
// per thread (8 waves per CU, up to 256 VGPRs each).
// For the pointwise stages each thread owns 4096 / threads representatives
// r = t + threads u, each standing for a bin K and its mirror -K (the spectra
// of real sequences satisfy Z[-K] = conj Z[K], so one value per
// representative is kept); the representatives cover the 8192 positions of
// the two tiles exactly once:
//   w >= 1:           K at tile 0 position r,   -K at tile 1 position F2-1-r;
//   w = 0, r < 2048:  K at tile 0 position r,   -K at tile 0 position F2-r,
//                     except r = 0, which holds the two self-mirrored bins
//                     0 and 2048 of row 0 (their spectra are real, so one
//                     double2 carries both: .x for bin 0, .y for bin 2048);
//   w = 0, r >= 2048: K at tile 1 position r-2048, -K at tile 1 F2-1-(r-2048).
// ---------------------------------------------------------------------------

// Row shape: 512 threads, tile = thread / 256, one radix-16 butterfly per
// thread and pass (three passes per 4096-point row, two LDS round trips),
// on the layout that keeps radix-16 passes free of bank conflicts.  KEEP:
// addresses held by the compiler (build, propagate) or derived afresh per
// pass (tangent build, whose four live spectra leave no registers for them).
template <bool K> struct RowR16 {
    static constexpr int NTH = 512, NT = 256, PTS = kTile;
    static constexpr bool KEEP = K;
    using TL = Swz16Layout<kF2>;
};
using RowBuild = RowR16<true>;
using RowProp = RowR16<true>;
using RowBuildT = RowR16<false>;
using RowPropFused = RowR16<false>;   // F = 8192 propagate: its loaders hold more state (33 spills with KEEP)
using RowBuildFused = RowR16<false>;  // F = 8192 build, likewise
using RowPropB = RowR16<false>;       // adjoint B propagate: four g spectra live

template <class TL> __host__ __device__ __forceinline__ int row_at(int k2) { return TL::at(0, k2); }

struct RowRep {
    int ia, ib;      // LDS positions of K and -K (r = 0 of w = 0: bins 0 and 2048)
    bool fixed;      // the two self-mirrored bins of row 0
};
template <class TL> __host__ __device__ __forceinline__ RowRep row_rep(int w, int r) {
    RowRep q;
    q.fixed = false;
    if (w != 0) {
        q.ia = row_at<TL>(r);
        q.ib = TL::IS + row_at<TL>(kF2 - 1 - r);
    } else if (r < kF2 / 2) {
        q.ia = row_at<TL>(r);
        q.ib = row_at<TL>(r == 0 ? kF2 / 2 : kF2 - r);
        q.fixed = (r == 0);
    } else {
        const int k2 = r - kF2 / 2;
        q.ia = TL::IS + row_at<TL>(k2);
        q.ib = TL::IS + row_at<TL>(kF2 - 1 - k2);
    }
    return q;
}
// Spectra X (real part's) and Y (imaginary part's) of the packed transform at
// a representative; for the fixed pair, (bin 0, bin 2048) of each, both real.
__device__ __forceinline__ void row_split(const RowRep &q, double2 a, double2 b, double2 &X, double2 &Y) {
    if (q.fixed) {
        X = make_double2(a.x, b.x);
        Y = make_double2(a.y, b.y);
    } else {
        X = split_x(a, b);
        Y = split_y(a, b);
    }
}
// Write A + iB at K and conj(A) + i conj(B) at -K (fixed pair: A, B real at
// both bins).
__device__ __forceinline__ void row_put(double2 *T, const RowRep &q, double2 A, double2 Bv) {
    if (q.fixed) {
        T[q.ia] = make_double2(A.x, Bv.x);
        T[q.ib] = make_double2(A.y, Bv.y);
    } else {
        T[q.ia] = pack_ai(A, Bv);
        T[q.ib] = pack_ai(c_conj(A), c_conj(Bv));
    }
}

// Rows of workgroup w: tile 0 and tile 1 (w = 0: the self-mirrored rows 0
// and F1/2; w > 0: the mirror pair w, F1 - w).  The row-pair kernels run
// rowpair_groups(F1) workgroups per (pair, q-plane).
__host__ __device__ __forceinline__ int row_of(int w, int F1, int h) {
    return (w == 0) ? (h ? F1 / 2 : 0) : (h ? F1 - w : w);
}
__host__ __device__ constexpr int rowpair_groups(int F1) { return F1 / 2; }

// Load both rows of one [k1][n2] array and transform them; the first radix-16
// pass reads its points (n2 = (t mod 256) + 256 s, coalesced) straight from global
// memory.  Ends with the last pass's barrier.
template <class RS>
__device__ __forceinline__ void rowpair_load_fft(double2 *T, const double2 *__restrict__ Zb, int rA, int rB,
                                                 const double2 *tw) {
    const double2 *z0 = Zb + (long long)rA * kF2, *z1 = Zb + (long long)rB * kF2;
    fft_tile_io<kF2, false, RS::NT, typename RS::TL, RS::PTS, RS::KEEP>(
        T, tw, [&](int h, int n2) { return (h ? z1 : z0)[n2]; }, NoIO{});
}

// Inverse-transform both tiles and store their rows of V (the inverse
// four-step twiddle is applied by the column inverse as it loads them); the
// last pass writes straight to global memory, and the tiles are free for
// writing on return.
template <class RS>
__device__ __forceinline__ void rowpair_ifft_store(double2 *T, double2 *__restrict__ Vb, int rA, int rB,
                                                   const double2 *tw) {
    fft_tile_io<kF2, true, RS::NT, typename RS::TL, RS::PTS, RS::KEEP>(
        T, tw, NoIO{}, [&](int h, int e, double2 v) { Vb[(long long)(h ? rB : rA) * kF2 + e] = v; });
}

// Where the row kernels get array a's rows and put its inverse rows.
// RowSrc / RowDst: the column kernels' [q][narr][F] arrays in HBM.
struct RowSrc {
    const double2 *base;
    int narr, logF;
    int qmul = 1;              // 0: one set of arrays shared by every q-plane (the payout's spectrum at the root)
    const int *map = nullptr;  // slot -> array (pruned propagates reading the per-parent spectrum cache)
    template <class RS>
    __device__ __forceinline__ void load_fft(double2 *T, int a, int qi, int rA, int rB, const double2 *tw) const {
        const int ai = map ? map[a] : a;
        rowpair_load_fft<RS>(T, base + (((long long)qi * qmul * narr + ai) << logF), rA, rB, tw);
    }
};
struct RowDst {
    double2 *base;
    int narr, logF;
    template <class RS>
    __device__ __forceinline__ void ifft_store(double2 *T, int a, int qi, int rA, int rB, const double2 *tw) const {
        rowpair_ifft_store<RS>(T, base + (((long long)qi * narr + a) << logF), rA, rB, tw);
    }
};
// F = 8192 (F1 = 2): workgroup 0 of a pair owns both rows, so the column
// stage runs here: rows k1 = 0, 1 of Y are x[n2] +- x[n2 + 4096], the second
// times w_F^n2, straight from the coefficient loader (Col2Src), and the
// inverse rows V0, V1 give x[n2] = V0 + V1 w_F^-n2 and x[n2 + 4096] =
// V0 - V1 w_F^-n2 for the coefficient storer (Col2Dst, which scales by 1/F).
// No column kernels and no column arrays for such a level.
// exp(-2 pi i j / 32) for a constant j (0 <= j < 32), correctly rounded.
__device__ __forceinline__ double2 w32(int j) {
    constexpr double c[32] = {
        1.0, 0.9807852804032304, 0.9238795325112867, 0.8314696123025452, 0.7071067811865476, 0.5555702330196022,
        0.3826834323650898, 0.19509032201612828, 0.0, -0.19509032201612828, -0.3826834323650898, -0.5555702330196022,
        -0.7071067811865476, -0.8314696123025452, -0.9238795325112867, -0.9807852804032304, -1.0, -0.9807852804032304,
        -0.9238795325112867, -0.8314696123025452, -0.7071067811865476, -0.5555702330196022, -0.3826834323650898,
        -0.19509032201612828, 0.0, 0.19509032201612828, 0.3826834323650898, 0.5555702330196022, 0.7071067811865476,
        0.8314696123025452, 0.9238795325112867, 0.9807852804032304};
    return make_double2(c[j & 31], -c[(j + 24) & 31]);   // sin(2 pi j / 32) = cos(2 pi (j - 8) / 32)
}
// The loader's points n2 = c + 256 s (s < 16) of one thread share c, so
// w_8192^n2 = w_8192^c w_32^s: one table twiddle per thread and a constant
// per point.  No table loads per point and no branch: a thread's loads then
// issue back to back (per-point twiddle loads under the tile's branch made
// every point wait for the previous one's loads).
template <class Rows> struct Col2Src {
    Rows src;
    int twShift;
    template <class RS>
    __device__ __forceinline__ void load_fft(double2 *T, int a, int qi, int, int, const double2 *tw) const {
        const RowPair r = src.rows(a, qi);
        fft_tile_io<kF2, false, RS::NT, typename RS::TL, RS::PTS, RS::KEEP>(
            T, tw,
            [&](int h, int n2) {
                const double2 x0 = r.at(n2), x1 = r.at(n2 + kF2);
                const double2 w = c_mul(tw_big2(tw, twShift, n2 & 255, 13), w32(n2 >> 8));
                const double2 d = c_mul(c_sub(x0, x1), w), sm = c_add(x0, x1);
                return h ? d : sm;
            },
            NoIO{});
    }
};
// The storer's n2 = t + 512 u: w_8192^n2 = w_8192^t w_32^(2u).
template <class St> struct Col2Dst {
    St st;
    int twShift;
    template <class RS>
    __device__ __forceinline__ void ifft_store(double2 *T, int a, int qi, int, int, const double2 *tw) const {
        using TL = typename RS::TL;
        static_assert(kF2 / RS::NTH <= 16, "w_32 steps");
        fft_tile_io<kF2, true, RS::NT, TL, RS::PTS, RS::KEEP>(T, tw, NoIO{}, NoIO{});
        const OutRows o = st.rows(a, qi);
        const double2 wt = tw_big2(tw, twShift, (int)threadIdx.x, 13);
#pragma unroll
        for (int u = 0; u < kF2 / RS::NTH; ++u) {
            const int n2 = (int)threadIdx.x + u * RS::NTH;
            const double2 v0 = T[TL::at(0, n2)];
            const double2 v1 = c_mulc(T[TL::IS + TL::at(0, n2)], c_mul(wt, w32(2 * u)));
            st.store(o, qi, n2, c_add(v0, v1));
            st.store(o, qi, n2 + kF2, c_sub(v0, v1));
        }
        __syncthreads();   // the tiles are free for writing on return
    }
};

// F = 8192 build level handing the level above (F' = 16384, F1' = 4) its
// child-pair column transforms: after the tile combine, coefficients n2 and
// n2 + 4096 of parents 2a, 2a+1 (scaled, exact coefficient 0, zero past the
// count) and the exact top term (index 8192, column 0) are column n2 of the
// level above; the 4-point column DFT and its twiddles w_16384^(n2 k1) are
// the operations k_col_fwd<4, ChildRows> applies to the stored coefficients
// (one radix-4 pass, then the same products), so the spectra are the same.
// Only each parent's two end coefficients are stored (for the exact terms
// of the level above).
struct Col2DstNext {
    LevelJob J;          // this level (F = 8192): parents, children, counts
    double2 *out;        // the level above's child-pair column transforms
    int nArr;            // parent pairs of this level = parents of the level above
    int twShift;
    __device__ __forceinline__ double2 coef(int n, double2 z, bool hx, bool hy, int qi, int p, double invF) const {
        double x = 0.0, y = 0.0;
        if (n == 0) {
            if (hx) x = child_prod_at(J, qi, p, 0);
            y = hy ? child_prod_at(J, qi, p + 1, 0) : 1.0;   // a missing right child is the identity
        } else if (n < J.parent.count) {
            if (hx) x = z.x * invF;
            if (hy) y = z.y * invF;
        }
        return make_double2(x, y);
    }
    template <class RS>
    __device__ __forceinline__ void ifft_store(double2 *T, int a, int qi, int, int, const double2 *tw) const {
        using TL = typename RS::TL;
        fft_tile_io<kF2, true, RS::NT, TL, RS::PTS, RS::KEEP>(T, tw, NoIO{}, NoIO{});
        const int p = 2 * a, pc = J.parent.count, F = J.F;
        const bool hx = p < J.nr_parent, hy = p + 1 < J.nr_parent;
        const double invF = 1.0 / (double)F;
        double *pb = J.parent.base + (long long)qi * J.parent.plane;
        double2 *o = out + ((long long)qi * nArr + a) * (2 * F);
        // twiddles of this level's storer (w_8192^t) and of the level
        // above's column transform (w_16384^(t r)), per thread
        const int t = (int)threadIdx.x;
        const double2 wt = tw_big2(tw, twShift, t, 13);
        const double2 wr1 = tw_big2(tw, twShift, t, 14), wr2 = tw_big2(tw, twShift, 2 * t, 14),
                      wr3 = tw_big2(tw, twShift, 3 * t, 14);
        double2 top = make_double2(0.0, 0.0);
        if (pc > F) {
            const int tc = J.child.count - 1;
            if (hx) top.x = child_prod_at(J, qi, p, tc);
            if (hy) top.y = child_prod_at(J, qi, p + 1, tc);
        }
#pragma unroll
        for (int u = 0; u < kF2 / RS::NTH; ++u) {
            const int n2 = (int)threadIdx.x + u * RS::NTH;
            const double2 v0 = T[TL::at(0, n2)];
            const double2 v1 = c_mulc(T[TL::IS + TL::at(0, n2)], c_mul(wt, w32(2 * u)));
            double2 a0 = coef(n2, c_add(v0, v1), hx, hy, qi, p, invF);
            double2 a1 = coef(n2 + kF2, c_sub(v0, v1), hx, hy, qi, p, invF);
            double2 a2 = n2 == 0 ? top : make_double2(0.0, 0.0);
            double2 a3 = make_double2(0.0, 0.0);
            // the end coefficients the level above reads for its exact terms
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int n = n2 + h * kF2;
                if (n == 0 || n == pc - 1) {
                    const double2 c = h ? a1 : a0;
                    if (hx) pb[(long long)p * J.parent.stride + n] = c.x;
                    if (hy) pb[(long long)(p + 1) * J.parent.stride + n] = c.y;
                }
            }
            if (n2 == 0 && pc > F) {
                if (hx) pb[(long long)p * J.parent.stride + F] = top.x;
                if (hy) pb[(long long)(p + 1) * J.parent.stride + F] = top.y;
            }
            dft4<false>(a0, a1, a2, a3);
            // w_16384^(n2 r) = w_16384^(t r) w_32^(u r)
            o[n2] = a0;
            o[kF2 + n2] = c_mul(a1, c_mul(wr1, w32(u)));
            o[2 * kF2 + n2] = c_mul(a2, c_mul(wr2, w32(2 * u)));
            o[3 * kF2 + n2] = c_mul(a3, c_mul(wr3, w32(3 * u)));
        }
        __syncthreads();   // the tiles are free for writing on return
    }
};

// Build: V[pair] rows = inverse rows of U = P_p + i P_{p+1}, P = X Y; rows
// of child pair (slot) p from src, inverse rows of parent pair `pair` to dst.
template <class RS, class Src, class Dst>
__global__ __launch_bounds__(RS::NTH) void k_rowpair_build(Src src, Dst dst, int nrp, int npairs, int logF,
                                                           const double2 *tw, int twShift) {
    constexpr int NTH = RS::NTH, RU = kTile / NTH;   // representatives per thread
    using TL = typename RS::TL;
    __shared__ double2 T[2 * kF2];
    const int F = 1 << logF, F1 = F >> 12;
    const int w = blockIdx.x, rA = row_of(w, F1, 0), rB = row_of(w, F1, 1);
    const int pair = blockIdx.y, qi = blockIdx.z;
    const int t = threadIdx.x;
    const int p = 2 * pair;
    (void)twShift;
    (void)npairs;
    double2 P1[RU], P2[RU];
    src.template load_fft<RS>(T, p, qi, rA, rB, tw);
#pragma unroll
    for (int u = 0; u < RU; ++u) {
        const RowRep q = row_rep<TL>(w, t + u * NTH);
        double2 X, Yv;
        row_split(q, T[q.ia], T[q.ib], X, Yv);
        P1[u] = q.fixed ? make_double2(X.x * Yv.x, X.y * Yv.y) : c_mul(X, Yv);
    }
    if (p + 1 < nrp) {  // uniform over the workgroup
        src.template load_fft<RS>(T, p + 1, qi, rA, rB, tw);
#pragma unroll
        for (int u = 0; u < RU; ++u) {
            const RowRep q = row_rep<TL>(w, t + u * NTH);
            double2 X, Yv;
            row_split(q, T[q.ia], T[q.ib], X, Yv);
            P2[u] = q.fixed ? make_double2(X.x * Yv.x, X.y * Yv.y) : c_mul(X, Yv);
        }
    } else {
#pragma unroll
        for (int u = 0; u < RU; ++u) P2[u] = make_double2(0.0, 0.0);
    }
    // every thread rewrites only the positions it read
#pragma unroll
    for (int u = 0; u < RU; ++u) row_put(T, row_rep<TL>(w, t + u * NTH), P1[u], P2[u]);
    dst.template ifft_store<RS>(T, pair, qi, rA, rB, tw);
}

// Propagate: V1[p] rows = inverse rows of U1 = g_p conj(R_p) + i g_p conj(L_p);
// child pairs (slot p) from srcc, g pairs (pair) from srcg, outputs (slot p)
// to dst.
template <class RS, class SrcC, class SrcG, class Dst>
__global__ __launch_bounds__(RS::NTH) void k_rowpair_prop(SrcC srcc, SrcG srcg, Dst dst, int nrp, int npairs,
                                                          int logF, const double2 *tw, int twShift) {
    constexpr int NTH = RS::NTH, RU = kTile / NTH;
    using TL = typename RS::TL;
    __shared__ double2 T[2 * kF2];
    const int F = 1 << logF, F1 = F >> 12;
    const int w = blockIdx.x, rA = row_of(w, F1, 0), rB = row_of(w, F1, 1);
    const int pair = blockIdx.y, qi = blockIdx.z;
    const int t = threadIdx.x;
    (void)twShift;
    (void)npairs;
    double2 G1[RU], G2[RU];
    srcg.template load_fft<RS>(T, pair, qi, rA, rB, tw);
#pragma unroll
    for (int u = 0; u < RU; ++u) {
        const RowRep q = row_rep<TL>(w, t + u * NTH);
        row_split(q, T[q.ia], T[q.ib], G1[u], G2[u]);
    }
    // one parent: its child pair's rows, U1 in place, inverse rows
    auto parent = [&](int p, const double2 (&Gw)[RU]) {
        srcc.template load_fft<RS>(T, p, qi, rA, rB, tw);
#pragma unroll
        for (int u = 0; u < RU; ++u) {
            const RowRep q = row_rep<TL>(w, t + u * NTH);
            double2 L, R;
            row_split(q, T[q.ia], T[q.ib], L, R);
            if (q.fixed)
                row_put(T, q, make_double2(Gw[u].x * R.x, Gw[u].y * R.y), make_double2(Gw[u].x * L.x, Gw[u].y * L.y));
            else
                row_put(T, q, c_mulc(Gw[u], R), c_mulc(Gw[u], L));   // g conj(R), g conj(L)
        }
        dst.template ifft_store<RS>(T, p, qi, rA, rB, tw);
    };
    parent(2 * pair, G1);
    if (2 * pair + 1 < nrp) parent(2 * pair + 1, G2);   // uniform over the workgroup
}

// Adjoint B propagate of a four-step level in one pass: per parent
//   U = A conj(dR) + B conj(R) + i (A conj(dL) + B conj(L))
// (A, B the parent's two adjoints; L + iR its children's products, dL + i dR
// their tangent products), one inverse per parent instead of one for each
// of the two correlations.  srca / srcb: the A and B g pairs, srcp / srcd:
// the child pairs' P and dP rows.
template <class RS, class SrcP, class SrcD, class SrcA, class SrcB, class Dst>
__global__ __launch_bounds__(RS::NTH) void k_rowpair_prop_b(SrcP srcp, SrcD srcd, SrcA srca, SrcB srcb, Dst dst,
                                                            int nrp, int npairs, int logF, const double2 *tw,
                                                            int twShift) {
    constexpr int NTH = RS::NTH, RU = kTile / NTH;
    using TL = typename RS::TL;
    __shared__ double2 T[2 * kF2];
    const int F = 1 << logF, F1 = F >> 12;
    const int w = blockIdx.x, rA = row_of(w, F1, 0), rB = row_of(w, F1, 1);
    const int pair = blockIdx.y, qi = blockIdx.z;
    const int t = threadIdx.x;
    (void)twShift;
    (void)npairs;
    double2 GA1[RU], GA2[RU], GB1[RU], GB2[RU];
    srca.template load_fft<RS>(T, pair, qi, rA, rB, tw);
#pragma unroll
    for (int u = 0; u < RU; ++u) {
        const RowRep q = row_rep<TL>(w, t + u * NTH);
        row_split(q, T[q.ia], T[q.ib], GA1[u], GA2[u]);
    }
    srcb.template load_fft<RS>(T, pair, qi, rA, rB, tw);
#pragma unroll
    for (int u = 0; u < RU; ++u) {
        const RowRep q = row_rep<TL>(w, t + u * NTH);
        row_split(q, T[q.ia], T[q.ib], GB1[u], GB2[u]);
    }
    auto parent = [&](int p, const double2 (&GA)[RU], const double2 (&GB)[RU]) {
        double2 UL[RU], UR[RU];
        srcd.template load_fft<RS>(T, p, qi, rA, rB, tw);
#pragma unroll
        for (int u = 0; u < RU; ++u) {
            const RowRep q = row_rep<TL>(w, t + u * NTH);
            double2 dL, dR;
            row_split(q, T[q.ia], T[q.ib], dL, dR);
            if (q.fixed) {
                UL[u] = make_double2(GA[u].x * dR.x, GA[u].y * dR.y);
                UR[u] = make_double2(GA[u].x * dL.x, GA[u].y * dL.y);
            } else {
                UL[u] = c_mulc(GA[u], dR);
                UR[u] = c_mulc(GA[u], dL);
            }
        }
        srcp.template load_fft<RS>(T, p, qi, rA, rB, tw);
#pragma unroll
        for (int u = 0; u < RU; ++u) {
            const RowRep q = row_rep<TL>(w, t + u * NTH);
            double2 L, R;
            row_split(q, T[q.ia], T[q.ib], L, R);
            if (q.fixed) {
                UL[u] = make_double2(UL[u].x + GB[u].x * R.x, UL[u].y + GB[u].y * R.y);
                UR[u] = make_double2(UR[u].x + GB[u].x * L.x, UR[u].y + GB[u].y * L.y);
            } else {
                UL[u] = c_add(UL[u], c_mulc(GB[u], R));
                UR[u] = c_add(UR[u], c_mulc(GB[u], L));
            }
            row_put(T, q, UL[u], UR[u]);
        }
        dst.template ifft_store<RS>(T, p, qi, rA, rB, tw);
    };
    parent(2 * pair, GA1, GB1);
    if (2 * pair + 1 < nrp) parent(2 * pair + 1, GA2, GB2);   // uniform over the workgroup
}

// Adjoint tangent build: V[pair] rows = inverse rows of
// U = dP_X[p] + i dP_X[p+1], dP_X = dP_L P_R + P_L dP_R, from the row-stage
// inputs of the children's P pairs (SP, the spectrum cache) and of their dP
// pairs (SD); both laid out [qi][nrp][F].
template <class RS>
__global__ __launch_bounds__(RS::NTH) void k_rowpair_build_t(const double2 *__restrict__ SP,
                                                         const double2 *__restrict__ SD, double2 *__restrict__ V,
                                                         int nrp, int npairs, int logF, const double2 *tw,
                                                         int twShift) {
    constexpr int NTH = RS::NTH, RU = kTile / NTH;
    using TL = typename RS::TL;
    __shared__ double2 T[2 * kF2];
    const int F = 1 << logF, F1 = F >> 12;
    const int w = blockIdx.x, rA = row_of(w, F1, 0), rB = row_of(w, F1, 1);
    const int pair = blockIdx.y, qi = blockIdx.z;
    const int t = threadIdx.x;
    (void)twShift;
    auto tangent = [&](int p, double2 (&D)[RU]) {
        rowpair_load_fft<RS>(T, SP + ((long long)qi * nrp + p) * F, rA, rB, tw);
        double2 PL[RU], PR[RU];
#pragma unroll
        for (int u = 0; u < RU; ++u) {
            const RowRep q = row_rep<TL>(w, t + u * NTH);
            row_split(q, T[q.ia], T[q.ib], PL[u], PR[u]);
        }
        rowpair_load_fft<RS>(T, SD + ((long long)qi * nrp + p) * F, rA, rB, tw);
#pragma unroll
        for (int u = 0; u < RU; ++u) {
            const RowRep q = row_rep<TL>(w, t + u * NTH);
            double2 DL, DR;
            row_split(q, T[q.ia], T[q.ib], DL, DR);
            D[u] = q.fixed ? make_double2(DL.x * PR[u].x + PL[u].x * DR.x, DL.y * PR[u].y + PL[u].y * DR.y)
                           : c_add(c_mul(DL, PR[u]), c_mul(PL[u], DR));
        }
    };
    double2 D1[RU], D2[RU];
    tangent(2 * pair, D1);
    if (2 * pair + 1 < nrp) {   // uniform over the workgroup
        tangent(2 * pair + 1, D2);
    } else {
#pragma unroll
        for (int u = 0; u < RU; ++u) D2[u] = make_double2(0.0, 0.0);
    }
    // every thread rewrites only the positions it read
#pragma unroll
    for (int u = 0; u < RU; ++u) row_put(T, row_rep<TL>(w, t + u * NTH), D1[u], D2[u]);
    rowpair_ifft_store<RS>(T, V + ((long long)qi * npairs + pair) * F, rA, rB, tw);
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
#define ICM_TRY(x)                              \
    do {                                        \
        hipError_t e_ = (x);                    \
        if (e_ != hipSuccess) return e_;        \
    } while (0)

static inline double fft_flops(double N) { return N <= 1 ? 0.0 : 5.0 * N * std::log2(N); }

// Profiler record names follow the kernel template instantiations.
static const char *kname(const char *fmt, int v) {
    static thread_local char buf[96];
    snprintf(buf, sizeof buf, fmt, v);
    return buf;
}

#define PF_BEGIN() do { if (pf && pf->on()) pf->begin(s); } while (0)
#define PF_END(name, bytes, flops) do { if (pf && pf->on()) pf->end(s, name, (bytes), (flops)); } while (0)

hipError_t launch_leaf_build(const LeafJob &J, hipStream_t s, Prof *pf) {
    dim3 grid((J.nb + 63) / 64, J.nq);
    PF_BEGIN();
    switch (J.B) {
    case 16: k_leaf_build<16><<<grid, 64, 0, s>>>(J); break;
    case 32: k_leaf_build<32><<<grid, 64, 0, s>>>(J); break;
    case 64: k_leaf_build<64><<<grid, 64, 0, s>>>(J); break;
    default: return hipErrorInvalidValue;
    }
    const double nq = J.nq, B = J.B;
    PF_END(kname("k_leaf_build<%d>", J.B), 8.0 * J.n + 8.0 * nq * J.nb * (B + 1), nq * J.nb * (1.5 * B * (B + 1) + 20.0 * B));
    return hipGetLastError();
}

hipError_t launch_leaf_extract(const LeafJob &J, int nchunk, hipStream_t s, Prof *pf) {
    const int bpw = 256 / J.B;
    dim3 grid((J.nbl + bpw - 1) / bpw, nchunk);
    PF_BEGIN();
    switch (J.B) {
    case 16: k_leaf_extract<16><<<grid, 256, 0, s>>>(J); break;
    case 32: k_leaf_extract<32><<<grid, 256, 0, s>>>(J); break;
    case 64: {
        const int ppl = leaf_ppl();
        if (ppl == 2) k_leaf_extract_n<64, 2><<<dim3((J.nbl + 7) / 8, nchunk), 256, 0, s>>>(J);
        else k_leaf_extract<64><<<grid, 256, 0, s>>>(J);
        break;
    }
    default: return hipErrorInvalidValue;
    }
    const double nq = J.nq, B = J.B;
    PF_END(kname("k_leaf_extract<%d>", J.B), 8.0 * nq * J.nbl * (B + 1 + J.G0) + 8.0 * J.n + 8.0 * nchunk * J.n,
           nq * (double)J.nbl * B * (4.0 * B + 30.0));
    return hipGetLastError();
}

hipError_t launch_leaf_build_vjp(const VjpLeafJob &V, hipStream_t s, Prof *pf) {
    const LeafJob &J = V.L;
    dim3 grid((J.nb + 63) / 64, J.nq);
    PF_BEGIN();
    switch (J.B) {
    case 16: k_leaf_build_vjp<16><<<grid, 64, 0, s>>>(V); break;
    case 32: k_leaf_build_vjp<32><<<grid, 64, 0, s>>>(V); break;
    case 64: k_leaf_build_vjp<64><<<grid, 64, 0, s>>>(V); break;
    default: return hipErrorInvalidValue;
    }
    const double nq = J.nq, B = J.B;
    PF_END(kname("k_leaf_build_vjp<%d>", J.B), 16.0 * J.n + 16.0 * nq * J.nb * (B + 1), nq * J.nb * 3.0 * B * (B + 1));
    return hipGetLastError();
}

hipError_t launch_leaf_vjp(const VjpLeafJob &V, int nchunk, hipStream_t s, Prof *pf) {
    const LeafJob &J = V.L;
    const int bpw = 256 / J.B;
    dim3 grid((J.nb + bpw - 1) / bpw, nchunk);
    PF_BEGIN();
    switch (J.B) {
    case 16: k_leaf_vjp<16><<<grid, 256, 0, s>>>(V); break;
    case 32: k_leaf_vjp<32><<<grid, 256, 0, s>>>(V); break;
    case 64:
        if (leaf_ppl() == 2) k_leaf_vjp_n<64, 2><<<dim3((J.nb + 7) / 8, nchunk), 256, 0, s>>>(V);
        else k_leaf_vjp<64><<<grid, 256, 0, s>>>(V);
        break;
    default: return hipErrorInvalidValue;
    }
    const double nq = J.nq, B = J.B;
    PF_END(kname("k_leaf_vjp<%d>", J.B), 8.0 * nq * J.nb * (3 * (B + 1) + V.c0) + 16.0 * J.n + 16.0 * nchunk * J.n,
           nq * J.n * (10.0 * B + 40.0));
    return hipGetLastError();
}

hipError_t launch_small(const SmallJob &J, int G, hipStream_t s, Prof *pf, double flops) {
    if (J.lds_doubles > kSmallLdsB || J.L > kSmallMaxL || G < 1) return hipErrorInvalidValue;
    PF_BEGIN();
    if (J.lds_doubles <= kSmallLdsA)
        k_small<kSmallLdsA><<<G, kSmallThreads, 0, s>>>(J);
    else
        k_small<kSmallLdsB><<<G, kSmallThreads, 0, s>>>(J);
    PF_END(J.lds_doubles <= kSmallLdsA ? "k_small<78KB>" : "k_small<152KB>",
           8.0 * ((double)J.n * 2 + 2.0 * J.Qa + J.k + (double)G * J.n), flops);
    return hipGetLastError();
}

hipError_t launch_small_finalize(const double *partial, int G, int n, const unsigned char *active, const int *perm,
                                 double *acc, double *equity, hipStream_t s, Prof *pf) {
    PF_BEGIN();
    const hipError_t e = colsum(partial, G, n, active, acc, 0, perm, equity, s);
    PF_END("k_small_finalize", 8.0 * n * ((double)G + 2.5), (double)n * G);
    return e;
}

hipError_t launch_stream_probe(int mode, const double *src, double *dst, long long n4, double *part, int nwg,
                               hipStream_t s) {
    if (mode == 3)
        k_lds_probe<<<nwg, 256, 0, s>>>(reinterpret_cast<const double2 *>(src), dst, (int)n4);
    else if (mode == 0)
        k_stream_read<<<nwg, 256, 0, s>>>(reinterpret_cast<const double4 *>(src), n4, part);
    else if (mode == 1)
        k_stream_write<<<nwg, 256, 0, s>>>(reinterpret_cast<double4 *>(dst), n4, 1.0);
    else
        k_stream_copy<<<nwg, 256, 0, s>>>(reinterpret_cast<const double4 *>(src), reinterpret_cast<double4 *>(dst),
                                          n4);
    return hipGetLastError();
}

hipError_t launch_gather(const double *src, const int *perm, int n, double *dst, hipStream_t s) {
    k_gather<<<(n + 255) / 256, 256, 0, s>>>(src, perm, n, dst);
    return hipGetLastError();
}

hipError_t launch_finalize(const double *partial, int nchunk, int n, const unsigned char *active, double *acc,
                           int accumulate, hipStream_t s, Prof *pf) {
    PF_BEGIN();
    const hipError_t e = colsum(partial, nchunk, n, active, acc, accumulate, nullptr, nullptr, s);
    PF_END("k_finalize", 8.0 * n * (nchunk + 2), (double)n * nchunk);
    return e;
}

hipError_t launch_scatter(const double *acc, int n, const int *perm, double *equity, hipStream_t s, Prof *pf) {
    PF_BEGIN();
    k_scatter<<<(n + 255) / 256, 256, 0, s>>>(acc, n, perm, equity);
    PF_END("k_scatter", 20.0 * n, 0.0);
    return hipGetLastError();
}

template <int F>
static hipError_t build_lds_f(const LevelJob &J, const double2 *tw, int twShift, hipStream_t s) {
    constexpr int ITEMS = kTile / F;
    const int npairs = (J.nr_parent + 1) / 2;
    dim3 grid((npairs + ITEMS - 1) / ITEMS, J.nq);
    k_build_lds<F><<<grid, lds_build_threads<F>(), 0, s>>>(J, tw, twShift);
    return hipGetLastError();
}
template <int F>
static hipError_t prop_lds_f(const LevelJob &J, const double2 *tw, int twShift, hipStream_t s) {
    constexpr int ITEMS = kTile / F;
    const int npairs = (J.np + 1) / 2;
    dim3 grid((npairs + ITEMS - 1) / ITEMS, J.nq);
    if (J.plist) k_prop_lds<F, true><<<grid, lds_prop_threads<F>(), 0, s>>>(J, tw, twShift);
    else k_prop_lds<F, false><<<grid, lds_prop_threads<F>(), 0, s>>>(J, tw, twShift);
    return hipGetLastError();
}

static double level_coef_bytes(const CoefView &v, int nodes, int nq) { return 8.0 * nq * (double)nodes * v.count; }

hipError_t launch_build_lds(const LevelJob &J, const double2 *tw, int twShift, hipStream_t s, Prof *pf) {
    PF_BEGIN();
    hipError_t e;
    switch (J.F) {
    case 16: e = build_lds_f<16>(J, tw, twShift, s); break;
    case 32: e = build_lds_f<32>(J, tw, twShift, s); break;
    case 64: e = build_lds_f<64>(J, tw, twShift, s); break;
    case 128: e = build_lds_f<128>(J, tw, twShift, s); break;
    case 256: e = build_lds_f<256>(J, tw, twShift, s); break;
    case 512: e = build_lds_f<512>(J, tw, twShift, s); break;
    case 1024: e = build_lds_f<1024>(J, tw, twShift, s); break;
    case 2048: e = build_lds_f<2048>(J, tw, twShift, s); break;
    case 4096: e = build_lds_f<4096>(J, tw, twShift, s); break;
    default: return hipErrorInvalidValue;
    }
    const double bytes = level_coef_bytes(J.child, J.nr_child, J.nq) + 8.0 * J.nq * (double)J.nr_parent * J.parent.count;
    PF_END(kname("k_build_lds<%d>", J.F), bytes, J.nq * (double)J.nr_parent * 1.5 * fft_flops(J.F));
    return e;
}

template <int F>
static hipError_t build_lds_t_f(const LevelJob &J, const double2 *tw, int twShift, hipStream_t s) {
    constexpr int ITEMS = kTile / F;
    dim3 grid((J.nr_parent + ITEMS - 1) / ITEMS, J.nq);
    k_build_lds_t<F><<<grid, lds_t_threads<F>(), 0, s>>>(J, tw, twShift);
    return hipGetLastError();
}
template <int F>
static hipError_t prop_lds_t_f(const LevelJob &J, const double2 *tw, int twShift, hipStream_t s) {
    constexpr int ITEMS = kTile / F;
    dim3 grid((J.nr_parent + ITEMS - 1) / ITEMS, J.nq);
    k_prop_lds_t<F><<<grid, lds_prop_t_threads<F>(), 0, s>>>(J, tw, twShift);
    return hipGetLastError();
}

#define ICM_FSWITCH(FN)                                              \
    switch (J.F) {                                                   \
    case 16: e = FN<16>(J, tw, twShift, s); break;                   \
    case 32: e = FN<32>(J, tw, twShift, s); break;                   \
    case 64: e = FN<64>(J, tw, twShift, s); break;                   \
    case 128: e = FN<128>(J, tw, twShift, s); break;                 \
    case 256: e = FN<256>(J, tw, twShift, s); break;                 \
    case 512: e = FN<512>(J, tw, twShift, s); break;                 \
    case 1024: e = FN<1024>(J, tw, twShift, s); break;               \
    case 2048: e = FN<2048>(J, tw, twShift, s); break;               \
    case 4096: e = FN<4096>(J, tw, twShift, s); break;               \
    default: return hipErrorInvalidValue;                            \
    }

hipError_t launch_build_lds_t(const LevelJob &J, const double2 *tw, int twShift, hipStream_t s, Prof *pf) {
    PF_BEGIN();
    hipError_t e;
    ICM_FSWITCH(build_lds_t_f)
    const double bytes = 2.0 * (level_coef_bytes(J.child, J.nr_child, J.nq) +
                                8.0 * J.nq * (double)J.nr_parent * J.parent.count);
    PF_END(kname("k_build_lds_t<%d>", J.F), bytes, J.nq * (double)J.nr_parent * 3.0 * fft_flops(J.F));
    return e;
}

hipError_t launch_prop_lds_t(const LevelJob &J, const double2 *tw, int twShift, hipStream_t s, Prof *pf) {
    PF_BEGIN();
    hipError_t e;
    ICM_FSWITCH(prop_lds_t_f)
    const double gin = (J.gpar.plane == 0 ? 8.0 * J.gpar.count : 8.0 * J.nq * (double)J.nr_parent * J.gpar.count) +
                       8.0 * J.nq * (double)J.nr_parent * J.bpar.count;
    const double bytes = gin + 2.0 * level_coef_bytes(J.child, J.nr_child, J.nq) +
                         8.0 * J.nq * (double)J.nr_child * (J.gchild.count + J.bchild.count);
    PF_END(kname("k_prop_lds_t<%d>", J.F), bytes, J.nq * (double)J.nr_parent * 5.0 * fft_flops(J.F));
    return e;
}

hipError_t launch_prop_lds(const LevelJob &J, const double2 *tw, int twShift, hipStream_t s, Prof *pf) {
    PF_BEGIN();
    hipError_t e;
    switch (J.F) {
    case 16: e = prop_lds_f<16>(J, tw, twShift, s); break;
    case 32: e = prop_lds_f<32>(J, tw, twShift, s); break;
    case 64: e = prop_lds_f<64>(J, tw, twShift, s); break;
    case 128: e = prop_lds_f<128>(J, tw, twShift, s); break;
    case 256: e = prop_lds_f<256>(J, tw, twShift, s); break;
    case 512: e = prop_lds_f<512>(J, tw, twShift, s); break;
    case 1024: e = prop_lds_f<1024>(J, tw, twShift, s); break;
    case 2048: e = prop_lds_f<2048>(J, tw, twShift, s); break;
    case 4096: e = prop_lds_f<4096>(J, tw, twShift, s); break;
    default: return hipErrorInvalidValue;
    }
    const double frac = (double)J.np / (double)J.nr_parent;   // < 1 with subset pruning
    const double gin = J.gpar.plane == 0 ? 8.0 * J.gpar.count : 8.0 * J.nq * (double)J.np * J.gpar.count;
    const double bytes = gin + frac * (level_coef_bytes(J.child, J.nr_child, J.nq) +
                                       8.0 * J.nq * (double)J.nr_child * J.gchild.count);
    PF_END(kname("k_prop_lds<%d>", J.F), bytes, J.nq * (double)J.np * 2.5 * fft_flops(J.F));
    return e;
}

template <class Rows>
static hipError_t col_fwd(int F1, Rows src, double2 *out, int nArr, int nq, int logF, const double2 *tw,
                          int twShift, hipStream_t s) {
    dim3 grid(F1, nArr, nq);  // tiles = F2 / TC = F1
#define ICM_CF(V) case V: k_col_fwd<V, Rows><<<grid, kCT, 0, s>>>(src, out, nArr, logF, tw, twShift); break;
    switch (F1) {
        ICM_CF(2) ICM_CF(4) ICM_CF(8) ICM_CF(16) ICM_CF(32) ICM_CF(64) ICM_CF(128) ICM_CF(256) ICM_CF(512)
        ICM_CF(1024) ICM_CF(2048) ICM_CF(4096)
    default: return hipErrorInvalidValue;
    }
#undef ICM_CF
    return hipGetLastError();
}

template <class Storer>
static hipError_t col_inv(int F1, const double2 *in, Storer st, int nArr, int nq, int logF, const double2 *tw,
                          int twShift, hipStream_t s) {
    dim3 grid(F1, nArr, nq);
#define ICM_CI(V) case V: k_col_inv<V, Storer><<<grid, kCT, 0, s>>>(in, st, nArr, logF, tw, twShift); break;
    switch (F1) {
        ICM_CI(2) ICM_CI(4) ICM_CI(8) ICM_CI(16) ICM_CI(32) ICM_CI(64) ICM_CI(128) ICM_CI(256) ICM_CI(512)
        ICM_CI(1024) ICM_CI(2048) ICM_CI(4096)
    default: return hipErrorInvalidValue;
    }
#undef ICM_CI
    return hipGetLastError();
}

static hipError_t col_inv_fwd_build(int F1, const double2 *in, const LevelJob &J, double2 *out, int nArr, int nq,
                                    int logF, const double2 *tw, int twShift, hipStream_t s, bool tan = false) {
    dim3 grid(F1, nArr, nq);
#define ICM_CIB(V)                                                                                            \
    case V:                                                                                                   \
        if (tan) k_col_inv_fwd_build<V, true><<<grid, kCT, 0, s>>>(in, J, out, nArr, logF, tw, twShift);     \
        else k_col_inv_fwd_build<V, false><<<grid, kCT, 0, s>>>(in, J, out, nArr, logF, tw, twShift);       \
        break;
    switch (F1) {
        ICM_CIB(4) ICM_CIB(8) ICM_CIB(16) ICM_CIB(32) ICM_CIB(64) ICM_CIB(128) ICM_CIB(256) ICM_CIB(512) ICM_CIB(1024)
    default: return hipErrorInvalidValue;
    }
#undef ICM_CIB
    return hipGetLastError();
}

bool build_fusable(const LevelJob &J, const LevelJob &Jn) {
    return (col_fused(J.F) || (J.F >= 4 * kF2 && J.F <= 1024 * kF2)) && Jn.F == 2 * J.F && !J.acc_out &&
           !J.child_r.base && !J.r_missing_zero && !Jn.child_r.base && !Jn.r_missing_zero &&
           Jn.nr_child == J.nr_parent && Jn.child.count == J.parent.count && J.parent.count <= J.F + 1 &&
           Jn.nq == J.nq;
}

hipError_t launch_build_large(const LevelJob &J, const double2 *tw, int twShift, double2 *bufA, double2 *bufB,
                              hipStream_t s, Prof *pf, double2 *spec, bool child_ready, const LevelJob *next,
                              double2 *next_spec) {
    if (next && (!next_spec || !build_fusable(J, *next))) return hipErrorInvalidValue;
    if (child_ready && !spec) return hipErrorInvalidValue;
    const int F1 = J.F >> 12;
    const int nrp = J.nr_parent, npairs = (nrp + 1) / 2;
    const double F = J.F, nq = J.nq;
    const double colFl = kF2 * fft_flops(F1) + 6.0 * F, rowFl = F1 * fft_flops(kF2);
    if (col_fused(J.F) && next) {   // column stage inside the row kernel, the level above's columns after it
        PF_BEGIN();
        k_rowpair_build<RowBuildFused, Col2Src<ChildRows>, Col2DstNext>
            <<<dim3(1, npairs, J.nq), RowBuildFused::NTH, 0, s>>>(Col2Src<ChildRows>{ChildRows{J}, twShift},
                                                            Col2DstNext{J, next_spec, npairs, twShift}, nrp,
                                                            npairs, J.logF, tw, twShift);
        ICM_TRY(hipGetLastError());
        PF_END("k_rowpair_build<F=8192,next>",
               level_coef_bytes(J.child, J.nr_child, J.nq) + 16.0 * nq * npairs * 2 * F,
               nq * (nrp + npairs) * (rowFl + colFl) + nq * npairs * 2 * F * 5.0);
        return hipSuccess;
    }
    if (col_fused(J.F)) {   // column stage inside the row kernel; nothing for the spectrum cache
        PF_BEGIN();
        k_rowpair_build<RowBuildFused, Col2Src<ChildRows>, Col2Dst<StParentPair>>
            <<<dim3(1, npairs, J.nq), RowBuildFused::NTH, 0, s>>>(Col2Src<ChildRows>{ChildRows{J}, twShift},
                                                            Col2Dst<StParentPair>{StParentPair{J}, twShift}, nrp,
                                                            npairs, J.logF, tw, twShift);
        ICM_TRY(hipGetLastError());
        PF_END("k_rowpair_build<F=8192>", level_coef_bytes(J.child, J.nr_child, J.nq) + 8.0 * nq * nrp * J.parent.count,
               nq * (nrp + npairs) * (rowFl + colFl));
        return hipSuccess;
    }
    double2 *Y = spec ? spec : bufA;
    if (!child_ready) {
        PF_BEGIN();
        ICM_TRY(col_fwd(F1, ChildRows{J}, Y, nrp, J.nq, J.logF, tw, twShift, s));
        PF_END(kname("k_col_fwd<%d,ChildRows>", F1),
               level_coef_bytes(J.child, J.nr_child, J.nq) + 16.0 * nq * nrp * F, nq * nrp * colFl);
    }
    PF_BEGIN();
    k_rowpair_build<RowBuild, RowSrc, RowDst><<<dim3(rowpair_groups(F1), npairs, J.nq), RowBuild::NTH, 0, s>>>(
        RowSrc{Y, nrp, J.logF}, RowDst{bufB, npairs, J.logF}, nrp, npairs, J.logF, tw, twShift);
    ICM_TRY(hipGetLastError());
    PF_END("k_rowpair_build", 16.0 * nq * (nrp + npairs) * F, nq * (nrp + npairs) * rowFl + 30.0 * nq * nrp * F);
    if (next) {   // the next level's child-pair spectra straight from this level's inverse columns
        PF_BEGIN();
        ICM_TRY(col_inv_fwd_build(F1, bufB, J, next_spec, npairs, J.nq, J.logF, tw, twShift, s));
        PF_END(kname("k_col_inv_fwd_build<%d>", F1), 16.0 * nq * npairs * F + 32.0 * nq * npairs * F,
               nq * npairs * kF2 * (fft_flops(F1) + fft_flops(2 * F1)));
        return hipSuccess;
    }
    PF_BEGIN();
    ICM_TRY(col_inv(F1, bufB, StParentPair{J}, npairs, J.nq, J.logF, tw, twShift, s));
    PF_END(kname("k_col_inv<%d,StParentPair>", F1), 16.0 * nq * npairs * F + 8.0 * nq * nrp * J.parent.count, nq * npairs * kF2 * fft_flops(F1));
    return hipSuccess;
}

hipError_t launch_build_large_t(const LevelJob &J, const double2 *tw, int twShift, const double2 *specP,
                                double2 *specD, double2 *bufB, hipStream_t s, Prof *pf, bool d_ready,
                                const LevelJob *next, double2 *next_specD) {
    if (!specP || !specD || col_fused(J.F)) return hipErrorInvalidValue;
    if (next && (!next_specD || !build_fusable_t(J, *next))) return hipErrorInvalidValue;
    const int F1 = J.F >> 12;
    const int nrp = J.nr_parent, npairs = (nrp + 1) / 2;
    const double F = J.F, nq = J.nq;
    const double colFl = kF2 * fft_flops(F1) + 6.0 * F, rowFl = F1 * fft_flops(kF2);
    LevelJob K = J;                     // the children's dP pairs, a missing right child 0
    K.child = J.tchild;
    K.child_r = CoefView{};
    K.r_missing_zero = 1;
    if (!d_ready) {
        PF_BEGIN();
        ICM_TRY(col_fwd(F1, ChildRows{K}, specD, nrp, J.nq, J.logF, tw, twShift, s));
        PF_END(kname("k_col_fwd<%d,ChildRows>", F1),
               level_coef_bytes(J.tchild, J.nr_child, J.nq) + 16.0 * nq * nrp * F, nq * nrp * colFl);
    }
    PF_BEGIN();
    k_rowpair_build_t<RowBuildT><<<dim3(rowpair_groups(F1), npairs, J.nq), RowBuildT::NTH, 0, s>>>(specP, specD, bufB, nrp, npairs,
                                                                                       J.logF, tw, twShift);
    ICM_TRY(hipGetLastError());
    PF_END("k_rowpair_build_t", 16.0 * nq * (2 * nrp + npairs) * F, nq * (2 * nrp + npairs) * rowFl + 40.0 * nq * nrp * F);
    if (next) {   // the level above's dP child-pair transforms straight from this level's inverse columns
        PF_BEGIN();
        ICM_TRY(col_inv_fwd_build(F1, bufB, J, next_specD, npairs, J.nq, J.logF, tw, twShift, s, true));
        PF_END(kname("k_col_inv_fwd_build_t<%d>", F1), 16.0 * nq * npairs * F + 32.0 * nq * npairs * F,
               nq * npairs * kF2 * (fft_flops(F1) + fft_flops(2 * F1)));
        return hipSuccess;
    }
    PF_BEGIN();
    ICM_TRY(col_inv(F1, bufB, StParentTangent{J}, npairs, J.nq, J.logF, tw, twShift, s));
    PF_END(kname("k_col_inv<%d,StParentTangent>", F1), 16.0 * nq * npairs * F + 8.0 * nq * nrp * J.parent.count,
           nq * npairs * kF2 * fft_flops(F1));
    return hipSuccess;
}

// The tangent hand-over: as build_fusable for the P pairs (same counts and
// transform sizes), with the tangent build's own composition (children's dP
// from tchild, nothing accumulated).
bool build_fusable_t(const LevelJob &J, const LevelJob &Jn) {
    return !col_fused(J.F) && J.F >= 4 * kF2 && J.F <= 1024 * kF2 && Jn.F == 2 * J.F && !J.acc_out &&
           !J.child_r.base && !J.r_missing_zero && J.tchild.base && Jn.nr_child == J.nr_parent &&
           Jn.child.count == J.parent.count && J.parent.count <= J.F + 1 && Jn.nq == J.nq;
}
static hipError_t col_inv_fwd(int F1, const double2 *in, const LevelJob &J, double2 *out, int nArr, int nq,
                              int logF, const double2 *tw, int twShift, hipStream_t s, bool keep_g0,
                              const LevelJob *J2 = nullptr) {
    dim3 grid(F1 / 2, nArr, nq);
#define ICM_CIF(V)                                                                                            \
    case V:                                                                                                   \
        if (J2) k_col_inv_fwd<V, true><<<grid, kCT, 0, s>>>(in, J, *J2, out, nArr, logF, tw, twShift, 0);    \
        else k_col_inv_fwd<V, false><<<grid, kCT, 0, s>>>(in, J, J, out, nArr, logF, tw, twShift, keep_g0 ? 1 : 0); \
        break;
    switch (F1) {
        ICM_CIF(8) ICM_CIF(16) ICM_CIF(32) ICM_CIF(64) ICM_CIF(128) ICM_CIF(256) ICM_CIF(512) ICM_CIF(1024)
        ICM_CIF(2048) ICM_CIF(4096)
    default: return hipErrorInvalidValue;
    }
#undef ICM_CIF
    return hipGetLastError();
}

bool prop_fusable(const LevelJob &J, const LevelJob &Jn) {
    return !col_fused(J.F) && J.F >= 8 * kF2 && Jn.F * 2 == J.F && !col_fused(Jn.F) && !J.plist && !Jn.plist &&
           !J.acc_out && !J.gfix && !J.child_r.base && !J.r_missing_zero && Jn.np == J.nr_child &&
           Jn.gpar.count == J.gchild.count && Jn.nq == J.nq && J.np == J.nr_parent;
}

hipError_t launch_g_pairs(const LevelJob &J, const double2 *tw, int twShift, double2 *bufA, hipStream_t s, Prof *pf) {
    if (J.F <= kMaxLdsF || col_fused(J.F)) return hipErrorInvalidValue;
    const int F1 = J.F >> 12;
    const int nrp = J.np, npairs = (nrp + 1) / 2;
    const long long r1 = (long long)J.nq * nrp * J.F;
    const int gq = J.gpar.plane == 0 ? 1 : J.nq;
    const double F = J.F, colFl = kF2 * fft_flops(F1) + 6.0 * F;
    const double gin = J.gpar.plane == 0 ? 8.0 * J.gpar.count : 8.0 * J.nq * (double)nrp * J.gpar.count;
    PF_BEGIN();
    ICM_TRY(col_fwd(F1, GRows{J}, bufA + r1, npairs, gq, J.logF, tw, twShift, s));
    PF_END(kname("k_col_fwd<%d,GRows>", F1), gin + 16.0 * gq * npairs * F, gq * npairs * colFl);
    return hipSuccess;
}

hipError_t launch_prop_large(const LevelJob &J, const double2 *tw, int twShift, double2 *bufA, double2 *bufB,
                             hipStream_t s, Prof *pf, const double2 *spec, bool g_ready, const LevelJob *next,
                             bool keep_g0) {
    if (next && !prop_fusable(J, *next)) return hipErrorInvalidValue;
    const int F1 = J.F >> 12;
    const int nrp = J.np, npairs = (nrp + 1) / 2;   // slots (all parents without pruning)
    const long long r1 = (long long)J.nq * nrp * J.F;   // parent region size
    const double F = J.F, nq = J.nq;
    const double colFl = kF2 * fft_flops(F1) + 6.0 * F, rowFl = F1 * fft_flops(kF2);
    const double gin = J.gpar.plane == 0 ? 8.0 * J.gpar.count : 8.0 * nq * nrp * J.gpar.count;
    if (col_fused(J.F)) {   // column stage inside the row kernel (spec and g_ready do not apply)
        PF_BEGIN();
        k_rowpair_prop<RowPropFused, Col2Src<ChildRows>, Col2Src<GRows>, Col2Dst<StGChildren>>
            <<<dim3(1, npairs, J.nq), RowPropFused::NTH, 0, s>>>(Col2Src<ChildRows>{ChildRows{J}, twShift},
                                                                Col2Src<GRows>{GRows{J}, twShift},
                                                           Col2Dst<StGChildren>{StGChildren{J}, twShift}, nrp, npairs,
                                                           J.logF, tw, twShift);
        ICM_TRY(hipGetLastError());
        PF_END("k_rowpair_prop<F=8192>",
               gin + level_coef_bytes(J.child, J.nr_child, J.nq) * nrp / J.nr_parent +
                   8.0 * nq * J.nr_child * J.gchild.count * nrp / J.nr_parent,
               nq * (2 * nrp + npairs) * (rowFl + colFl));
        return hipSuccess;
    }
    // cached spectra are per parent: pruned launches (slots over a parent
    // list) read them through the list
    const double2 *Y1 = spec ? spec : bufA;
    const bool mapped = spec && J.plist;
    if (!spec) {
        PF_BEGIN();
        ICM_TRY(col_fwd(F1, ChildRows{J}, bufA, nrp, J.nq, J.logF, tw, twShift, s));
        PF_END(kname("k_col_fwd<%d,ChildRows>", F1),
               level_coef_bytes(J.child, J.nr_child, J.nq) + 16.0 * nq * nrp * F, nq * nrp * colFl);
    }
    // the root's g is the payout, the same for every q-plane: its pair
    // spectrum is computed once and read by all planes
    const int gq = J.gpar.plane == 0 ? 1 : J.nq;
    if (!g_ready) {
        PF_BEGIN();
        ICM_TRY(col_fwd(F1, GRows{J}, bufA + r1, npairs, gq, J.logF, tw, twShift, s));
        PF_END(kname("k_col_fwd<%d,GRows>", F1), gin + 16.0 * gq * npairs * F, gq * npairs * colFl);
    }
    PF_BEGIN();
    k_rowpair_prop<RowProp, RowSrc, RowSrc, RowDst><<<dim3(rowpair_groups(F1), npairs, J.nq), RowProp::NTH, 0, s>>>(
        mapped ? RowSrc{Y1, J.nr_parent, J.logF, 1, J.plist} : RowSrc{Y1, nrp, J.logF},
        RowSrc{bufA + r1, npairs, J.logF, gq == 1 && J.nq > 1 ? 0 : 1}, RowDst{bufB, nrp, J.logF}, nrp, npairs,
        J.logF, tw, twShift);
    ICM_TRY(hipGetLastError());
    PF_END("k_rowpair_prop", 16.0 * nq * (2 * nrp + npairs) * F, nq * (2 * nrp + npairs) * rowFl + 40.0 * nq * nrp * F);
    if (next) {   // the next level's g pairs straight from this level's inverse columns
        double2 *gout = bufA + (long long)next->nq * next->np * next->F;
        PF_BEGIN();
        ICM_TRY(col_inv_fwd(F1, bufB, J, gout, nrp, J.nq, J.logF, tw, twShift, s, keep_g0));
        PF_END(kname("k_col_inv_fwd<%d>", F1), 16.0 * nq * nrp * F + 8.0 * nq * nrp * F,
               nq * nrp * kF2 * (fft_flops(F1) + fft_flops(F1 / 2)));
        return hipSuccess;
    }
    PF_BEGIN();
    ICM_TRY(col_inv(F1, bufB, StGChildren{J}, nrp, J.nq, J.logF, tw, twShift, s));
    PF_END(kname("k_col_inv<%d,StGChildren>", F1), 16.0 * nq * nrp * F + 8.0 * nq * J.nr_child * J.gchild.count, nq * nrp * kF2 * fft_flops(F1));
    return hipSuccess;
}

bool prop_b_fusable(const LevelJob &JdR, const LevelJob &Jn) {
    return !col_fused(JdR.F) && JdR.F >= 8 * kF2 && Jn.F * 2 == JdR.F && !col_fused(Jn.F) && !JdR.plist &&
           !Jn.plist && !JdR.acc_out && Jn.np == JdR.nr_child && Jn.gpar.count == JdR.gchild.count &&
           Jn.nq == JdR.nq && JdR.np == JdR.nr_parent;
}

hipError_t launch_prop_large_b(const LevelJob &JdR, const LevelJob &JR, const double2 *tw, int twShift,
                               double2 *bufA, double2 *bufB, const double2 *specP, const double2 *specD, hipStream_t s,
                               Prof *pf, bool b_ready, const LevelJob *next) {
    if (col_fused(JdR.F)) {   // F = 8192: the column stage inside the row kernel, straight from the coefficients
        if (JdR.plist || b_ready || next) return hipErrorInvalidValue;
        const int nrp = JdR.np, npairs = (nrp + 1) / 2;
        const double F = JdR.F, nq = JdR.nq;
        const double colFl = kF2 * fft_flops(2) + 6.0 * F, rowFl = 2 * fft_flops(kF2);
        PF_BEGIN();
        k_rowpair_prop_b<RowPropB, Col2Src<ChildRows>, Col2Src<ChildRows>, Col2Src<GRows>, Col2Src<GRows>,
                         Col2Dst<StGChildren2>><<<dim3(1, npairs, JdR.nq), RowPropB::NTH, 0, s>>>(
            Col2Src<ChildRows>{ChildRows{JR}, twShift}, Col2Src<ChildRows>{ChildRows{JdR}, twShift},
            Col2Src<GRows>{GRows{JdR}, twShift}, Col2Src<GRows>{GRows{JR}, twShift},
            Col2Dst<StGChildren2>{StGChildren2{JdR, JR}, twShift}, nrp, npairs, JdR.logF, tw, twShift);
        ICM_TRY(hipGetLastError());
        PF_END("k_rowpair_prop_b<F=8192>",
               8.0 * nq * ((double)nrp * (JdR.gpar.count + JR.gpar.count) + 2.0 * JdR.nr_child * JdR.child.count) +
                   8.0 * nq * JdR.nr_child * JdR.gchild.count,
               nq * (4 * nrp + 2 * npairs) * (rowFl + colFl));
        return hipSuccess;
    }
    if (!specP || !specD || JdR.plist) return hipErrorInvalidValue;
    if (next && !prop_b_fusable(JdR, *next)) return hipErrorInvalidValue;
    const int F1 = JdR.F >> 12;
    const int nrp = JdR.np, npairs = (nrp + 1) / 2;
    const long long r1 = (long long)JdR.nq * nrp * JdR.F;
    const double F = JdR.F, nq = JdR.nq;
    const double colFl = kF2 * fft_flops(F1) + 6.0 * F, rowFl = F1 * fft_flops(kF2);
    const double gin = JR.gpar.plane == 0 ? 8.0 * JR.gpar.count : 8.0 * nq * nrp * JR.gpar.count;
    if (!b_ready) {
        PF_BEGIN();   // the B pairs into the child-pair region of bufA (unused with the caches)
        ICM_TRY(col_fwd(F1, GRows{JR}, bufA, npairs, JdR.nq, JdR.logF, tw, twShift, s));
        PF_END(kname("k_col_fwd<%d,GRows>", F1), gin + 16.0 * nq * npairs * F, nq * npairs * colFl);
    }
    PF_BEGIN();
    k_rowpair_prop_b<RowPropB, RowSrc, RowSrc, RowSrc, RowSrc, RowDst><<<dim3(rowpair_groups(F1), npairs, JdR.nq), RowPropB::NTH, 0, s>>>(
        RowSrc{specP, nrp, JdR.logF}, RowSrc{specD, nrp, JdR.logF}, RowSrc{bufA + r1, npairs, JdR.logF},
        RowSrc{bufA, npairs, JdR.logF}, RowDst{bufB, nrp, JdR.logF}, nrp, npairs, JdR.logF, tw, twShift);
    ICM_TRY(hipGetLastError());
    PF_END("k_rowpair_prop_b", 16.0 * nq * (3 * nrp + 2 * npairs) * F, nq * (3 * nrp + 2 * npairs) * rowFl);
    if (next) {   // the level below's B pairs straight from this level's inverse columns
        PF_BEGIN();
        ICM_TRY(col_inv_fwd(F1, bufB, JdR, bufA, nrp, JdR.nq, JdR.logF, tw, twShift, s, false, &JR));
        PF_END(kname("k_col_inv_fwd_b<%d>", F1), 16.0 * nq * nrp * F + 8.0 * nq * nrp * F,
               nq * nrp * kF2 * (fft_flops(F1) + fft_flops(F1 / 2)));
        return hipSuccess;
    }
    PF_BEGIN();
    ICM_TRY(col_inv(F1, bufB, StGChildren2{JdR, JR}, nrp, JdR.nq, JdR.logF, tw, twShift, s));
    PF_END(kname("k_col_inv<%d,StGChildren2>", F1),
           16.0 * nq * nrp * F + 8.0 * nq * JdR.nr_child * JdR.gchild.count, nq * nrp * kF2 * fft_flops(F1));
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// Ragged-batch launchers
// ---------------------------------------------------------------------------
// Leaf build of the batch at B = 64: eight lanes per block (94 VGPRs, four
// waves per SIMD) unless ICM_HIP_LEAF_COOP=0 (read at each launch; a captured
// graph keeps the choice of its capture) selects one lane per block.
static bool leaf_coop() {
    const char *e = getenv("ICM_HIP_LEAF_COOP");
    return !(e && atoi(e) == 0);
}

hipError_t launch_leaf_build_batch(int B, const BatchInst *inst, const int *blk_inst, const double *S,
                                   double *blocks, int NB, int nq, int q0, hipStream_t s, Prof *pf, double bytes) {
    if (NB < 1 || nq < 1) return hipSuccess;
    PF_BEGIN();
    if (B == 64 && leaf_coop()) {
        // about 4096 workgroups: columns of 64 blocks, rows of q-planes
        const int cols = (NB + 63) / 64;
        const int rows = std::max(1, std::min(nq, (4096 + cols - 1) / cols));
        const int qper = (nq + rows - 1) / rows;
        k_leaf_build_batch_coop<64><<<dim3(cols, (nq + qper - 1) / qper), kLT, 0, s>>>(inst, blk_inst, S, blocks, NB,
                                                                                      nq, q0, qper);
        hipError_t e = hipGetLastError();
        PF_END("k_leaf_build_batch_coop", bytes, 0.0);
        return e;
    }
    dim3 grid((NB + 63) / 64, nq);
    switch (B) {
    case 16: k_leaf_build_batch<16><<<grid, 64, 0, s>>>(inst, blk_inst, S, blocks, NB, q0); break;
    case 32: k_leaf_build_batch<32><<<grid, 64, 0, s>>>(inst, blk_inst, S, blocks, NB, q0); break;
    case 64: k_leaf_build_batch<64><<<grid, 64, 0, s>>>(inst, blk_inst, S, blocks, NB, q0); break;
    default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    PF_END("k_leaf_build_batch", bytes, 0.0);
    return e;
}

hipError_t launch_leaf_extract_batch(int B, const BatchInst *inst, const int *blk_inst, const double *S,
                                     const double *blocks, int NB, int nq, int q0, int nchunk, int q_per_chunk,
                                     double *partial, int Ntot, hipStream_t s, Prof *pf, double bytes) {
    if (NB < 1 || nq < 1) return hipSuccess;
    PF_BEGIN();
    const int bpw = 256 / B;
    dim3 grid((NB + bpw - 1) / bpw, nchunk);
    switch (B) {
    case 16:
        k_leaf_extract_batch<16><<<grid, 256, 0, s>>>(inst, blk_inst, S, blocks, NB, nq, q0, q_per_chunk, partial,
                                                     Ntot);
        break;
    case 32:
        k_leaf_extract_batch<32><<<grid, 256, 0, s>>>(inst, blk_inst, S, blocks, NB, nq, q0, q_per_chunk, partial,
                                                     Ntot);
        break;
    case 64: {
        const int ppl = leaf_ppl();
        if (ppl == 2)
            k_leaf_extract_batch_n<64, 2><<<dim3((NB + 7) / 8, nchunk), 256, 0, s>>>(
                inst, blk_inst, S, blocks, NB, nq, q0, q_per_chunk, partial, Ntot);
        else
            k_leaf_extract_batch<64><<<grid, 256, 0, s>>>(inst, blk_inst, S, blocks, NB, nq, q0, q_per_chunk,
                                                         partial, Ntot);
        break;
    }
    default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    PF_END("k_leaf_extract_batch", bytes, 0.0);
    return e;
}

hipError_t launch_level_batch(bool build, const LevelJob *jobs, const int2 *items, int nitems, const double2 *tw,
                              int twShift, hipStream_t s, Prof *pf, double bytes, double flops) {
    if (nitems < 1) return hipSuccess;
    PF_BEGIN();
    if (build) k_build_lds_batch<<<nitems, kLT, 0, s>>>(jobs, items, tw, twShift);
    else k_prop_lds_batch<<<nitems, kLT, 0, s>>>(jobs, items, tw, twShift);
    hipError_t e = hipGetLastError();
    PF_END(build ? "k_build_lds_batch" : "k_prop_lds_batch", bytes, flops);
    return e;
}

hipError_t launch_leaf_build_vjp_batch(int B, const BatchInst *inst, const int *blk_inst, const double *S,
                                       const double *v, double *blocks, double *dblocks, int NB, int nq, int q0,
                                       hipStream_t s, Prof *pf, double bytes) {
    if (NB < 1 || nq < 1) return hipSuccess;
    PF_BEGIN();
    dim3 grid((NB + 63) / 64, nq);
    switch (B) {
    case 16: k_leaf_build_vjp_batch<16><<<grid, 64, 0, s>>>(inst, blk_inst, S, v, blocks, dblocks, NB, q0); break;
    case 32: k_leaf_build_vjp_batch<32><<<grid, 64, 0, s>>>(inst, blk_inst, S, v, blocks, dblocks, NB, q0); break;
    case 64: k_leaf_build_vjp_batch<64><<<grid, 64, 0, s>>>(inst, blk_inst, S, v, blocks, dblocks, NB, q0); break;
    default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    PF_END("k_leaf_build_vjp_batch", bytes, 0.0);
    return e;
}

hipError_t launch_leaf_vjp_batch(int B, const BatchInst *inst, const int *blk_inst, const double *S,
                                 const double *v, const double *blocks, const double *dblocks, int NB, int nq,
                                 int q0, int nchunk, int q_per_chunk, double *partial, double *partial_ds, int Ntot,
                                 hipStream_t s, Prof *pf, double bytes) {
    if (NB < 1 || nq < 1) return hipSuccess;
    PF_BEGIN();
    const int bpw = 256 / B;
    dim3 grid((NB + bpw - 1) / bpw, nchunk);
#define ICM_LVB(BB)                                                                                          \
    k_leaf_vjp_batch<BB><<<grid, 256, 0, s>>>(inst, blk_inst, S, v, blocks, dblocks, NB, nq, q0, q_per_chunk, \
                                              partial, partial_ds, Ntot)
    switch (B) {
    case 16: ICM_LVB(16); break;
    case 32: ICM_LVB(32); break;
    case 64: ICM_LVB(64); break;
    default: return hipErrorInvalidValue;
    }
#undef ICM_LVB
    hipError_t e = hipGetLastError();
    PF_END("k_leaf_vjp_batch", bytes, 0.0);
    return e;
}

hipError_t launch_level_t_batch(bool build, const LevelJob *jobs, const int2 *items, int nitems, const double2 *tw,
                                int twShift, hipStream_t s, Prof *pf, double bytes, double flops) {
    if (nitems < 1) return hipSuccess;
    PF_BEGIN();
    if (build) k_build_lds_t_batch<<<nitems, kLT, 0, s>>>(jobs, items, tw, twShift);
    else k_prop_lds_t_batch<<<nitems, kLT, 0, s>>>(jobs, items, tw, twShift);
    hipError_t e = hipGetLastError();
    PF_END(build ? "k_build_lds_t_batch" : "k_prop_lds_t_batch", bytes, flops);
    return e;
}

hipError_t launch_build_sb(const LevelJob &J, hipStream_t s, Prof *pf) {
    if (J.child.count > kSbMaxC || sb_stride_build(J.child.count) > kSbLds) return hipErrorInvalidValue;
    PF_BEGIN();
    const int ppw = sb_ppw(sb_stride_build(J.child.count));
    dim3 grid((J.nr_parent + ppw - 1) / ppw, J.nq);
    k_build_sb<<<grid, kSbT, 0, s>>>(J, ppw);
    hipError_t e = hipGetLastError();
    const double bytes = level_coef_bytes(J.child, J.nr_child, J.nq) + 8.0 * J.nq * (double)J.nr_parent * J.parent.count;
    PF_END(kname("k_build_sb<c%d>", J.child.count), bytes,
           2.0 * J.nq * (double)J.nr_parent * J.child.count * J.parent.count);
    return e;
}

hipError_t launch_prop_sb(const LevelJob &J, hipStream_t s, Prof *pf) {
    if (J.child.count > kSbMaxC || sb_stride_prop(J.child.count, J.gchild.count) > kSbLds)
        return hipErrorInvalidValue;
    PF_BEGIN();
    const int ppw = sb_ppw(sb_stride_prop(J.child.count, J.gchild.count));
    dim3 grid((J.np + ppw - 1) / ppw, J.nq);
    k_prop_sb<<<grid, kSbT, 0, s>>>(J, ppw);
    hipError_t e = hipGetLastError();
    const double gin = (J.gpar.plane == 0 ? 8.0 * J.gpar.count : 8.0 * J.nq * (double)J.np * J.gpar.count);
    const double bytes = gin + 8.0 * J.nq * (double)J.np * 2.0 * (J.child.count + J.gchild.count);
    PF_END(kname("k_prop_sb<c%d>", J.child.count), bytes,
           4.0 * J.nq * (double)J.np * J.child.count * J.gchild.count);
    return e;
}

hipError_t launch_level_sb_batch(bool build, const LevelJob *jobs, const int2 *items, int nitems, hipStream_t s,
                                 Prof *pf, double bytes, double flops) {
    if (nitems < 1) return hipSuccess;
    PF_BEGIN();
    if (build) k_build_sb_batch<<<nitems, kSbT, 0, s>>>(jobs, items);
    else k_prop_sb_batch<<<nitems, kSbT, 0, s>>>(jobs, items);
    hipError_t e = hipGetLastError();
    PF_END(build ? "k_build_sb_batch" : "k_prop_sb_batch", bytes, flops);
    return e;
}

// The persistent kernel is built for the batch's tuned block size, 64; other
// block sizes run the per-level launches.
int persist_extract_blocks(int B) { return B == 64 ? PersistCfg<64>::EXB : 0; }

template <int B> static int persist_grid_b() {
    int dev = 0, cus = 0, per = 0;
    if (hipFuncSetAttribute(reinterpret_cast<const void *>(k_batch_persist<B>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, kPersistLds) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_batch_persist<B>, kLT, kPersistLds) != hipSuccess)
        return 0;
    return cus * per;
}

int batch_persist_grid(int B) { return B == 64 ? persist_grid_b<64>() : 0; }

hipError_t launch_batch_persist(int B, const PersistArgs &A, int grid, hipStream_t s, Prof *pf, double bytes,
                                double flops) {
    if (A.nitems < 1) return hipSuccess;
    if (grid < 1) return hipErrorInvalidValue;
    PF_BEGIN();
    if (B != 64) return hipErrorInvalidValue;
    k_batch_persist<64><<<grid, kLT, kPersistLds, s>>>(A);
    hipError_t e = hipGetLastError();
    PF_END("k_batch_persist", bytes, flops);
    return e;
}


// Host check of the index maps the row-pair kernels are built on, run by the
// library's self-check entry point (no device needed):
//  * every LDS layout maps (transform, index) one to one onto its tile;
//  * for every F1 = F / 4096 up to 4096, the rows of the rowpair_groups(F1)
//    workgroups cover the F1 rows exactly once;
//  * in every workgroup, the representatives r = 0 .. 4095 of row_rep cover
//    the 2 x 4096 LDS positions of its two tiles exactly once, and each pair
//    (ia, ib) holds mirror bins of the F-point spectrum: with bin
//    K = row + F1 k2, bin(ib) = F - bin(ia) mod F (both self-mirrored for the
//    fixed pair).  A wrong mirror or a position written twice or never would
//    mix or drop spectra in the row-pair kernels.
template <class TL> static bool layout_bijective(std::string &why) {
    std::vector<char> seen(TL::SIZE, 0);
    for (int it = 0; it < TL::ITEMS; ++it) {
        for (int idx = 0; idx < kTile / TL::ITEMS; ++idx) {
            const int a = TL::at(it, idx);
            if (a < 0 || a >= TL::SIZE || seen[a]++) {
                why = "layout position out of range or repeated";
                return false;
            }
        }
    }
    return true;
}

bool selfcheck_row_layouts(std::string &why) {
    bool ok = layout_bijective<SwzLayout<16>>(why) && layout_bijective<SwzLayout<32>>(why) &&
              layout_bijective<SwzLayout<64>>(why) && layout_bijective<SwzLayout<128>>(why) &&
              layout_bijective<SwzLayout<1024>>(why) && layout_bijective<SwzLayout<4096>>(why) &&
              layout_bijective<Swz16Layout<128>>(why) && layout_bijective<Swz16Layout<256>>(why) &&
              layout_bijective<Swz16Layout<512>>(why) && layout_bijective<Swz16Layout<1024>>(why) &&
              layout_bijective<Swz16Layout<2048>>(why) && layout_bijective<Swz16Layout<4096>>(why) &&
              layout_bijective<ColLayout<16>>(why) && layout_bijective<ColLayout<512>>(why) &&
              layout_bijective<ColLayout<4096>>(why);
    if (!ok) return false;
    using TL = RowR16<true>::TL;
    std::vector<int> inv(TL::IS, -1);
    for (int k2 = 0; k2 < kF2; ++k2) inv[row_at<TL>(k2)] = k2;
    for (int F1 = 2; F1 <= 4096; F1 *= 2) {
        const long long F = (long long)F1 * kF2;
        std::vector<char> row_seen(F1, 0);
        for (int w = 0; w < rowpair_groups(F1); ++w) {
            const int rows[2] = {row_of(w, F1, 0), row_of(w, F1, 1)};
            for (int h = 0; h < 2; ++h) {
                if (rows[h] < 0 || rows[h] >= F1 || row_seen[rows[h]]++) {
                    why = "row-pair workgroups do not cover the rows exactly once at F1 = " + std::to_string(F1);
                    return false;
                }
            }
            std::vector<char> seen(2 * TL::IS, 0);
            auto bin = [&](int pos) {
                const int h = pos / TL::IS;
                return (long long)rows[h] + (long long)F1 * inv[pos % TL::IS];
            };
            for (int r = 0; r < kF2; ++r) {
                const RowRep q = row_rep<TL>(w, r);
                if (q.ia < 0 || q.ib < 0 || q.ia >= 2 * TL::IS || q.ib >= 2 * TL::IS || seen[q.ia]++ || seen[q.ib]++) {
                    why = "row representatives do not cover the two tiles exactly once at F1 = " +
                          std::to_string(F1) + ", w = " + std::to_string(w);
                    return false;
                }
                const long long ka = bin(q.ia), kb = bin(q.ib);
                const bool mirror = q.fixed ? ((F - ka) % F == ka && (F - kb) % F == kb && ka != kb)
                                            : ((F - ka) % F == kb);
                if (!mirror) {
                    why = "row representative pairs non-mirror bins at F1 = " + std::to_string(F1) +
                          ", w = " + std::to_string(w) + ", r = " + std::to_string(r);
                    return false;
                }
            }
        }
    }
    return true;
}

}  // namespace icm
