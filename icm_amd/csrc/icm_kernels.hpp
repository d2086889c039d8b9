// icm_kernels.hpp -- launch interface of the HIP kernels (host side).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace icm {

// Device view of one tree level for the build / propagate kernels.
// Coefficients of node j of q-plane qi live at
//   base + qi * plane + j * stride + m,  m < count.
struct CoefView {
    const double *base;
    long long plane;
    int stride;
    int count;
};
struct CoefOut {
    double *base;
    long long plane;
    int stride;
    int count;
};

struct LevelJob {
    CoefView child;     // level l-1 products (left children; right too unless child_r)
    CoefOut parent;     // level l products (build)
    CoefView gpar;      // level l g-vectors (propagate); plane 0 at the root
    CoefOut gchild;     // level l-1 g-vectors (propagate)
    int nr_child;       // real nodes at level l-1
    int nr_parent;      // real nodes at level l
    int F;              // transform size of this level
    int logF;
    int nq;             // q-planes in the launch
    // Propagate launches may run over a list of parents (subset pruning):
    // slot i handles parent plist[i] (i < np); without a list, parent i
    // (np = nr_parent).  Build launches always use the full level.
    const int *plist;
    int np;
    // Composition for the adjoint pass (zero / null for the forward pass):
    CoefView child_r;   // right children from another array (base == nullptr: child)
    int r_missing_zero; // a missing right child is 0 instead of the identity
    int acc_out;        // add to the outputs instead of storing
    int gfix;           // propagate outputs keep c_child terms with F = 2c - 2:
                        // add the one wrapped term X[c-1] (Y[F] - Y[0])
    // Fused adjoint kernels (k_build_lds_t / k_prop_lds_t):
    CoefView tchild;    // tangent products of the children
    CoefOut tparent;    // tangent product of the parent (build)
    CoefView bpar;      // B adjoint of the parent (propagate; base null = 0 at the root)
    CoefOut bchild;     // B adjoint of the children (propagate)
};

struct LeafJob {
    const double *S;        // sorted stacks [n]
    const double *logv;     // [nq]
    const double *wq;       // [nq] quadrature weights
    const double *a_in;     // optional [nq][n] a values (replaces exp(S logv))
    double *blocks;         // [nq][nb][B+1]
    const double *g;        // g at the leaves: base, plane, stride
    long long g_plane;
    int g_stride;
    int G0;                 // g entries per block
    int n, nb, B, k;
    int nq;
    double *partial;        // [nchunk][n] equity partials (sorted order)
    double *inner_out;      // optional [nq][n] Psi values (sorted order)
    int q_per_chunk;
    const unsigned char *active;  // optional [n] (sorted order)
    const int *blist;       // optional list of blocks to extract (subset pruning)
    int nbl;                // blocks in the launch (= nb without a list)
};

// Adjoint (VJP) leaf data: tangent block products and the two adjoints.
struct VjpLeafJob {
    LeafJob L;              // blocks = P_b, g = A_b (forward g), partial = equity partials
    const double *v;        // cotangent, sorted order [n]
    const double *c_in;     // optional [nq][n] tangent weights (replaces v * pw)
    double *dblocks;        // [nq][nb][B+1] tangent block products
    const double *bvec;     // B adjoint at the leaves: base, plane, stride (null: zero)
    long long b_plane;
    int b_stride;
    int c0;                 // B entries per block (= level-0 coefficient count)
    double *partial_ds;     // [nchunk][n] dS partials (sorted order)
    double *h_out;          // optional [nq][n] H_j values
};

// One instance of a ragged batch (device table, one entry per instance).
struct BatchInst {
    const double *logv;     // nodes of the instance, indexed by absolute q
    const double *w;
    CoefView g0;            // g at the leaves (the payout, plane 0, when L == 1)
    int n, k, nb, L;
    int poff;               // first player in the concatenated sorted stacks
    int boff;               // first block in the global block numbering
    int G0;                 // g entries per block
    CoefView b0;            // adjoint pass: B adjoint at the leaves (base null when L == 1: zero)
    int c0;                 // B entries per block (= level-0 coefficient count)
};

// Optional per-launch timing: when on, every kernel launch is bracketed by
// two events on the launch stream and tagged with its compulsory HBM bytes
// and FLOPs (inputs read once + outputs written once; 5 N log2 N per complex
// transform of size N).
struct Prof {
    virtual ~Prof() {}
    virtual bool on() const = 0;
    virtual void begin(hipStream_t s) = 0;
    virtual void end(hipStream_t s, const char *name, double bytes, double flops) = 0;
};

// Leaf (block) build: block products of B players per (block, q).
hipError_t launch_leaf_build(const LeafJob &J, hipStream_t s, Prof *pf);
// Leaf divide + dot + quadrature-weighted accumulation.
hipError_t launch_leaf_extract(const LeafJob &J, int nchunk, hipStream_t s, Prof *pf);
// acc[j] (+)= sum_c partial[c][j] (fixed order), masked for subsets.
hipError_t launch_finalize(const double *partial, int nchunk, int n, const unsigned char *active,
                           double *acc, int accumulate, hipStream_t s, Prof *pf);
hipError_t launch_scatter(const double *acc, int n, const int *perm, double *equity, hipStream_t s,
                          Prof *pf);

// Adjoint leaves: (P_b, dP_b) per block; then Psi, H and dS per player.
hipError_t launch_leaf_build_vjp(const VjpLeafJob &J, hipStream_t s, Prof *pf);
hipError_t launch_leaf_vjp(const VjpLeafJob &J, int nchunk, hipStream_t s, Prof *pf);
hipError_t launch_gather(const double *src, const int *perm, int n, double *dst, hipStream_t s);
// Streaming probes over n4 double4 units with nwg workgroups of 256 threads:
// mode 0 reads src (one partial sum per workgroup into part[nwg]), 1 fills
// dst, 2 copies src to dst; mode 3 is the LDS read probe (src: 2048 double2,
// dst: nwg * 256 doubles, n4 = reads per lane).
hipError_t launch_stream_probe(int mode, const double *src, double *dst, long long n4, double *part, int nwg,
                               hipStream_t s);

// Fused adjoint levels (F <= 4096): P and dP together; A and B together.
hipError_t launch_build_lds_t(const LevelJob &J, const double2 *tw, int twShift, hipStream_t s, Prof *pf);
hipError_t launch_prop_lds_t(const LevelJob &J, const double2 *tw, int twShift, hipStream_t s, Prof *pf);

// One tree level, LDS-resident transforms (F <= 4096).
hipError_t launch_build_lds(const LevelJob &J, const double2 *tw, int twShift, hipStream_t s, Prof *pf);
hipError_t launch_prop_lds(const LevelJob &J, const double2 *tw, int twShift, hipStream_t s, Prof *pf);

// One tree level through HBM with a three-pass four-step transform (F >= 8192).
// bufA / bufB: scratch of (n_real[l] + ceil(n_real[l]/2)) * F * nq complex each.
// At F = 8192 (col_fused) the column stage runs inside the row kernels, which
// read and write the coefficients directly: bufA, bufB and spec are unused.
// spec (optional, nq * n_real[l] * F complex): the build keeps the column
// transforms of the child pairs there, and the propagate of the same level
// and q-planes reads them instead of transforming the children again (only
// without a parent list: slots are then the parents themselves).
// child_ready: spec already holds this level's child-pair column transforms
// (written by the level below's fused column inverse).  next / next_spec
// (optional): the build of the level above (F * 2, a four-step level with
// column kernels) is fused in: this level's column inverse writes next's
// child-pair column transforms into next_spec instead of this level's
// parent coefficients, which are then never stored (only where
// build_fusable, and where every later reader of them takes the spectra).
hipError_t launch_build_large(const LevelJob &J, const double2 *tw, int twShift, double2 *bufA,
                              double2 *bufB, hipStream_t s, Prof *pf, double2 *spec = nullptr,
                              bool child_ready = false, const LevelJob *next = nullptr,
                              double2 *next_spec = nullptr);
bool build_fusable(const LevelJob &J, const LevelJob &next);
// Adjoint tangent build of a large level: dP_X = dP_L P_R + P_L dP_R into
// J.parent (J.child = P, J.tchild = dP of the children).  specP holds the
// level's P-pair row-stage inputs (the spectrum cache, written by the P
// build); the dP-pair inputs are written to specD (nq * n_real[l] * F
// complex) for the A * dR propagate of the same q-batch to reuse.
// d_ready: specD already holds this level's dP-pair transforms (the level
// below handed them over).  next / next_specD (optional): the tangent build
// of the level above is fused in: this level's column inverse writes its dP
// child-pair transforms into next_specD; of this level's dP only the end
// coefficients (0, count - 1 and the top term) are stored, for the exact
// terms that read them (only where build_fusable_t, and where every later
// reader takes the spectra).
hipError_t launch_build_large_t(const LevelJob &J, const double2 *tw, int twShift, const double2 *specP,
                                double2 *specD, double2 *bufB, hipStream_t s, Prof *pf, bool d_ready = false,
                                const LevelJob *next = nullptr, double2 *next_specD = nullptr);
bool build_fusable_t(const LevelJob &J, const LevelJob &next);
// g_ready: bufA already holds this J's g-pair column transforms (the previous
// launch had the same gpar, nq and parents), so they are not computed again.
// next (optional): the propagate of the level below (F / 2, also a four-step
// level with column kernels) is fused in: this level's column inverse writes
// next's g-pair column transforms (at bufA + next.nq * next.np * next.F,
// where next's launch reads them with g_ready) instead of the children's g
// coefficients, which are then never stored (keep_g0: except each child's
// coefficient 0).  Only where prop_fusable.
hipError_t launch_prop_large(const LevelJob &J, const double2 *tw, int twShift, double2 *bufA,
                             double2 *bufB, hipStream_t s, Prof *pf, const double2 *spec = nullptr,
                             bool g_ready = false, const LevelJob *next = nullptr, bool keep_g0 = false);
// The g-pair column transforms of a four-step propagate J into bufA (where
// launch_prop_large with g_ready and launch_prop_large_b read them).
hipError_t launch_g_pairs(const LevelJob &J, const double2 *tw, int twShift, double2 *bufA, hipStream_t s,
                          Prof *pf);
bool prop_fusable(const LevelJob &J, const LevelJob &next);

// Adjoint B propagate of a four-step level in one pass: B children =
// A * dR + B * R (JdR: A's g pair over the dP children, JR: B's g pair over
// the P children; JdR's outputs and flags).  Needs both spectrum caches
// (specP: the P pairs' column transforms, specD: the dP pairs') and bufA
// holding JdR's g-pair column transforms (the A propagate just ran); B's
// pairs go to bufA's child-pair region.  No parent lists.
// b_ready: bufA's child-pair region already holds B's g-pair transforms (the
// level above handed them over).  next (optional): the level below's B
// propagate (its JR) is fused in: this level's column inverse writes its B
// g-pair transforms into that region; of the B children only coefficients 0
// and F / 2 are stored, for the wrapped terms (only where prop_b_fusable).
hipError_t launch_prop_large_b(const LevelJob &JdR, const LevelJob &JR, const double2 *tw, int twShift,
                               double2 *bufA, double2 *bufB, const double2 *specP, const double2 *specD, hipStream_t s,
                               Prof *pf, bool b_ready = false, const LevelJob *next = nullptr);
bool prop_b_fusable(const LevelJob &JdR, const LevelJob &next);

// Ragged batches.  Leaves: global blocks [q][NB][B+1] for q-planes
// q0 .. q0 + nq - 1; extract writes partial[chunk][Ntot].  Levels: one
// launch per tree level over nitems work items (job, first tile item);
// bytes / flops are the launch's compulsory totals, for the profiler.
hipError_t launch_leaf_build_batch(int B, const BatchInst *inst, const int *blk_inst, const double *S,
                                   double *blocks, int NB, int nq, int q0, hipStream_t s, Prof *pf, double bytes);
hipError_t launch_leaf_extract_batch(int B, const BatchInst *inst, const int *blk_inst, const double *S,
                                     const double *blocks, int NB, int nq, int q0, int nchunk, int q_per_chunk,
                                     double *partial, int Ntot, hipStream_t s, Prof *pf, double bytes);
hipError_t launch_level_batch(bool build, const LevelJob *jobs, const int2 *items, int nitems, const double2 *tw,
                              int twShift, hipStream_t s, Prof *pf, double bytes, double flops);
// Ragged-batch adjoint: (P, dP) leaves, fused adjoint levels (one parent per
// tile item), leaf adjoint with equity and dS partials.  v is sorted order.
hipError_t launch_leaf_build_vjp_batch(int B, const BatchInst *inst, const int *blk_inst, const double *S,
                                       const double *v, double *blocks, double *dblocks, int NB, int nq, int q0,
                                       hipStream_t s, Prof *pf, double bytes);
hipError_t launch_leaf_vjp_batch(int B, const BatchInst *inst, const int *blk_inst, const double *S,
                                 const double *v, const double *blocks, const double *dblocks, int NB, int nq,
                                 int q0, int nchunk, int q_per_chunk, double *partial, double *partial_ds, int Ntot,
                                 hipStream_t s, Prof *pf, double bytes);
hipError_t launch_level_t_batch(bool build, const LevelJob *jobs, const int2 *items, int nitems, const double2 *tw,
                                int twShift, hipStream_t s, Prof *pf, double bytes, double flops);

constexpr int kMaxLdsF = 4096;
// Four-step levels whose column stage (F1 = 2) runs inside the row kernels:
// no column arrays and no spectrum-cache entry.
__host__ __device__ inline bool col_fused(int F) { return F == 2 * kMaxLdsF; }

// Schoolbook levels: direct FP64 sums for levels whose children keep few
// coefficients.  A workgroup stages the operands of `ppw` parents in
// kSbLds doubles of LDS; the counts below are shared by host and device.
constexpr int kSbLds = 4096;
constexpr int kSbMaxC = 512;    // largest child coefficient count the kernels take
__host__ __device__ inline int sb_stride_build(int cc) { return 2 * cc + 10; }
__host__ __device__ inline int sb_stride_prop(int cc, int gc) { return 2 * cc + gc + cc + 8; }
__host__ __device__ inline int sb_ppw(int stride) {
    const int p = kSbLds / stride;
    return p < 1 ? 1 : (p > 64 ? 64 : p);
}

// One tree level by schoolbook sums (children count c_{l-1} <= kSbMaxC).
hipError_t launch_build_sb(const LevelJob &J, hipStream_t s, Prof *pf);
hipError_t launch_prop_sb(const LevelJob &J, hipStream_t s, Prof *pf);
hipError_t launch_level_sb_batch(bool build, const LevelJob *jobs, const int2 *items, int nitems, hipStream_t s,
                                 Prof *pf, double bytes, double flops);

// Persistent ragged batch: one launch per q-pass runs every phase (leaf
// build, build and propagate levels, extract, finalize) as ticketed lists of
// work items with per-instance completion counters (no barrier between
// levels), one list per XCD: instance ranges [qinst[x], qinst[x+1]) are
// served only by workgroups on XCD x, so an instance's data stays in one L2.
// Item kinds and fields:
//   kPLeaf      a = first global block (kPLeafBlocks blocks, every q-plane)
//   kPBuildLds, kPBuildSb, kPPropLds, kPPropSb
//               a = job (LevelJob index), b = first tile item (as the
//               per-level launches), need = finished items of the instance
//               required first
//   kPExtract   a = first global block (persist_extract_blocks(B)), b = q-chunk
//   kPFinal     a = first player (sorted order, kPFinalPlayers players)
// kind_inst = kind | (first instance << 3); leaf, extract and final items
// span instances first .. need (the need counts of extract and final come
// from need_ex / need_fin per instance).
enum PersistKind { kPLeaf = 0, kPBuildLds = 1, kPBuildSb = 2, kPPropLds = 3, kPPropSb = 4, kPExtract = 5, kPFinal = 6 };
struct PItem {
    int kind_inst;
    int a;
    int b;
    int need;
};
constexpr int kPLeafBlocks = 64;
constexpr int kPFinalPlayers = 512;
// Queues (one per XCD of MI355X) and the scheduler array: sched[x] ticket of
// queue x, sched[kPSchedDone] workgroups finished, sched[kPSchedErr] error
// (1: a wait passed spin_limit, 2: a queue no workgroup served).
constexpr int kPMaxQueues = 8;
constexpr int kPSchedDone = kPMaxQueues;
constexpr int kPSchedErr = kPMaxQueues + 1;
constexpr int kPSchedSize = 16;
struct PersistArgs {
    const PItem *items;       // queue x: items[qoff[x] .. qoff[x+1])
    int nitems;
    const int *qoff;          // [nqueues + 1]
    const int *qend;          // [nqueues][2]: end of queue x's global blocks, of its players
    int nqueues;
    const LevelJob *jobs;
    const BatchInst *inst;
    const int *blk_inst;
    const double *S;          // sorted stacks, concatenated
    double *blocks;           // [nq][NB][B+1]
    int NB, nq, q0, ninst;
    const double2 *tw;
    int twShift;
    int nchunk, q_per_chunk;
    double *partial;          // [nchunk][Ntot]
    int Ntot;
    double *acc;              // [Ntot] running total over passes (sorted order)
    int accumulate;           // add acc (passes after the first)
    const int *perm;          // sorted position -> flat original index
    double *equity;           // last pass: equity[perm[j]] = acc[j]; null otherwise
    const int *need_ex;       // [ninst] finished items before an instance's extract
    const int *need_fin;      // [ninst] finished items before its finalize
    int *ctr;                 // [ninst] finished items (zero at launch; cleared by the last workgroup)
    int *sched;               // [kPSchedSize], layout above
    int spin_limit;
};
// Blocks per extract item of the persistent kernel (two groups of 256
// threads); 0 for a block size the kernel is not built for (only 64 is).
int persist_extract_blocks(int B);
// Workgroups of the persistent kernel resident at once on the current device
// (0 on error); the launch uses at most this many.
int batch_persist_grid(int B);
hipError_t launch_batch_persist(int B, const PersistArgs &A, int grid, hipStream_t s, Prof *pf, double bytes,
                                double flops);

// Small instances in one launch: one workgroup per quadrature node (a
// workgroup loops over nodes b, b + G, ... when there are more nodes than
// workgroups), the whole tree with blocks of one player in LDS (product
// levels 1 .. L-2, two g levels), schoolbook sums per level, Psi_j = g_j[0]
// at the leaves, and the weighted contribution added to the workgroup's own
// row of partial[G][n] in node order (deterministic).  Geometry: B = 1,
// n_real[l] = ceil(n_real[l-1] / 2), c[l] = min(2^l + 1, k),
// g[l] = min(2^l, k); off[l]: LDS offset (doubles) of level l's products
// (l = 1 .. L-2); a_j at offset 0; the g levels at goff[0 / 1].
constexpr int kSmallMaxL = 16;
constexpr int kSmallThreads = 1024;
struct SmallJob {
    const double *S;        // sorted stacks [n]
    const double *logv;     // [Qa]
    const double *w;        // [Qa]
    const double *payout;   // [k]
    double *partial;        // [G][n] (sorted order)
    int n, k, L, Qa;
    int nr[kSmallMaxL], c[kSmallMaxL], g[kSmallMaxL], off[kSmallMaxL];
    int goff[2];
    int lds_doubles;        // LDS the geometry needs
};
// LDS tiers of the small-instance kernel (doubles): two workgroups per CU,
// or one workgroup with most of the CU's 160 KB.
constexpr int kSmallLdsA = 10000;
constexpr int kSmallLdsB = 19456;
// G workgroups; partial rows are overwritten (not accumulated).
hipError_t launch_small(const SmallJob &J, int G, hipStream_t s, Prof *pf, double flops);
// equity[perm[j]] = sum_{b < G} partial[b][j] in order (0 for inactive j).
hipError_t launch_small_finalize(const double *partial, int G, int n, const unsigned char *active, const int *perm,
                                 double *acc, double *equity, hipStream_t s, Prof *pf);

// Host check of the row-pair kernels' index maps (no device needed); false
// with the reason in why on the first violation.
bool selfcheck_row_layouts(std::string &why);

}  // namespace icm
