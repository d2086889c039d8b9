/*
 * icm_hip.h -- C ABI of the MI355X ICM equity engine (libicm_hip.so).
 *
 * Drop-in boundary for the reference's two public headers:
 *   - src/cpu/icm.h:26-41        icm_init / icm_equity / icm_equity_subset
 *   - src/gpu/icm_gpu.h:10-116   IcmGpuPlan, options, stats, plan API
 * Plain pointers and sizes only.  All host pointers belong to the caller;
 * the library copies what it keeps (S at plan creation, payout per call).
 *
 * Return conventions follow the reference: the icm_hip_* / icm_gpu_* compute
 * calls return 0 on success and -1 on failure (message in
 * icm_hip_last_error()); icm_hip_init / icm_gpu_init return 1 on success and
 * 0 on failure (src/gpu/gpu_api.cu:61-81); the CPU-named aliases return void
 * like src/cpu/icm.h and leave the output zeroed on failure.
 */
#ifndef ICM_HIP_H
#define ICM_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct IcmHipPlan IcmHipPlan;

/* Field order and types mirror IcmGpuOptions (src/gpu/icm_gpu.h:12-34) so a
 * caller built against the reference header passes a compatible struct.
 * Fields keep the reference meaning where this engine has the feature and
 * are otherwise accepted at their defaults and ignored; a value this engine
 * cannot honour is rejected at plan creation (NULL, icm_hip_last_error()). */
typedef struct {
    int device_id;                    /* informational; set by the library */
    int use_cufftdx;                  /* ignored (no vendor FFT library is used) */
    int enable_graphs;                /* ragged batches: record/replay a HIP graph; ignored by single plans */
    int enable_q_pipeline;            /* ignored: all q-planes of a batch run in one launch per level */
    int memory_strategy;              /* 3 = selective recompute: no spectrum cache (the propagate
                                       * transforms the children again); any other value: device memory
                                       * per plan, cache on when it fits (src/gpu/gpu_plan.cu:335) */
    int force_uncached_fused_levels;  /* -1 or 0 = auto; this engine has no per-level cache count, so
                                       * positive values are rejected */
    int force_uncached_cufft_levels;  /* -1 or 0 = auto; positive values are rejected (as above) */
    int fast_mode;                    /* ignored */
    int verbose;                      /* print the plan to stderr */
} IcmHipOptions;

/* Plan parameters that have no IcmGpuOptions field (no reference
 * counterpart; the reference reads ICM_GPU_Q_BATCH, src/gpu/gpu_api.cu:139-144,
 * and sizes from a fixed VRAM share, gpu_api.cu:187).  Zero = default.
 * Environment: ICM_GPU_Q_BATCH sets q_batch when the field is 0;
 * ICM_HIP_MEM_BUDGET (bytes) sets mem_budget_bytes when the field is 0;
 * ICM_HIP_MEMORY_STRATEGY overrides IcmHipOptions.memory_strategy (also for
 * the one-shot entries, which take no options). */
typedef struct {
    int block_size;                   /* 16, 32 or 64; 0 = MI355X tuning table */
    int q_batch;                      /* q-planes per launch (<= 256); 0 = from the memory budget */
    double mem_fraction;              /* share of free device memory a plan plans for; 0 = 0.6 */
    size_t mem_budget_bytes;          /* cap on the plan's device bytes: planning uses
                                       * min(free, cap) and an allocation beyond the cap fails as if
                                       * the device had run out; 0 = no cap */
} IcmHipPlanParams;

/* Mirrors IcmGpuRunStats (src/gpu/icm_gpu.h:36-50). */
typedef struct {
    double total_ns;                  /* device time of the run (events on the plan's stream) */
    double sort_ns;
    double quadrature_overhead_ns;
    double block_build_ns;
    double tree_build_ns;
    double tree_propagate_cached_ns;
    double tree_propagate_recomputed_ns;
    double leaf_extract_ns;
    size_t peak_vram_bytes;
    int engine;                       /* 1 = hybrid tree */
    int B;
    int uncached_fused_levels;        /* number of LDS-resident levels */
    int uncached_cufft_levels;        /* number of four-step (HBM) levels */
} IcmHipRunStats;

/* Mirrors IcmGpuPlanSummary (src/gpu/icm_gpu.h:52-63). */
typedef struct {
    int n;
    int k;
    int B;
    int engine;
    int n_levels;
    int n_tier1;                      /* levels with transform size <= 64 */
    int n_tier2;                      /* other LDS-resident levels (<= 4096) */
    int n_tier3;                      /* four-step levels (>= 8192) */
    int q_batch;
    size_t planned_peak_vram_bytes;
} IcmHipPlanSummary;

/* ---- lifecycle (src/gpu/icm_gpu.h:66-76) ---- */
int icm_hip_init(int device_id);               /* 1 = ok, 0 = failure */
void icm_hip_shutdown(void);
const char *icm_hip_last_error(void);
/* Trims the device's default stream-ordered memory pool to 0 (plans
 * allocate with hipMalloc and free on destroy, so this only returns what
 * other hipMallocAsync users left pooled).  0 = ok, -1 before icm_hip_init or
 * when the trim fails (src/gpu/icm_gpu.h:70-74, gpu_api.cu:94-102). */
int icm_hip_release_pooled_memory(void);

/* ---- plan (src/gpu/icm_gpu.h:79-81) ----
 * Device memory is allocated once per plan, sized from mem_fraction of free
 * memory; when an allocation fails the q-batch is halved and the plan
 * allocated again, up to 4 times (src/gpu/gpu_memory.cu:267-279). */
IcmHipPlan *icm_hip_plan_create(int n, const double *S, int k, const IcmHipOptions *opts);
/* As icm_hip_plan_create with IcmHipPlanParams (params may be NULL). */
IcmHipPlan *icm_hip_plan_create_ex(int n, const double *S, int k, const IcmHipOptions *opts,
                                   const IcmHipPlanParams *params);
/* Host-only planning (no device call): runs the planner and the allocation
 * sequence with its OOM retries against a simulated device whose free memory
 * is params->mem_budget_bytes (required).  Fills the summary the plan would
 * have and the number of q-batch halvings; returns 0, or -1 when even one
 * q-plane does not fit (message in icm_hip_last_error()). */
int icm_hip_plan_dry_run(int n, int k, const IcmHipOptions *opts, const IcmHipPlanParams *params,
                         IcmHipPlanSummary *summary, int *retries);
void icm_hip_plan_destroy(IcmHipPlan *plan);
int icm_hip_plan_summary(const IcmHipPlan *plan, IcmHipPlanSummary *summary);

/* ---- compute (src/gpu/icm_gpu.h:94-111) ---- */
int icm_hip_equity_with_plan(IcmHipPlan *plan, int Q, const double *payout, double *equity,
                             IcmHipRunStats *stats);
int icm_hip_equity(int n, const double *S, int Q, const double *payout, int k, double *equity);
int icm_hip_equity_subset(int n, const double *S, int Q, const double *payout, int k,
                          double *equity, const int *targets, int n_targets);

/* Device-resident variant of icm_hip_equity_with_plan: d_payout[k] and
 * d_equity[n] are device pointers; nothing crosses PCIe.  Runs on the plan's
 * stream and returns after the stream has finished.  No reference
 * counterpart (the reference copies payout in and equity out per call,
 * src/gpu/gpu_api.cu:270,455). */
int icm_hip_equity_with_plan_device(IcmHipPlan *plan, int Q, const double *d_payout,
                                    double *d_equity, IcmHipRunStats *stats);

/* Restrict later runs of a plan to target players (0-based, original order;
 * out-of-range entries ignored): non-targets come out 0 and the propagate /
 * leaf kernels run only over subtrees holding a target (the pruning of
 * src/cpu/icm.c:2430-2521).  targets == NULL or n_targets <= 0 restores all
 * players.  No reference counterpart (the reference's subset call builds a
 * fresh plan and masks only the outputs, src/gpu/gpu_api.cu:620-672). */
int icm_hip_plan_set_targets(IcmHipPlan *plan, const int *targets, int n_targets);

/* One quadrature node through the GPU tree: inner[i] = Psi_i for the given
 * a[i] (players in the given order), mirroring the IcmEngine hook
 * (src/cpu/icm.h:45-47, engine_hybrid_ctx src/cpu/icm.c:2307-2318). */
int icm_hip_node_inner(int n, const double *a, const double *payout, int k, double *inner);

/* ---- adjoint (no reference counterpart: the reference has no gradient) ----
 * dS[i] = d(v . equity)/dS_i for a cotangent v[n] (original player order),
 * with the quadrature nodes held fixed (they depend on max S only through
 * y_hi).  The pass also returns the equities (either output may be NULL).
 * Plans restricted with icm_hip_plan_set_targets are rejected. */
int icm_hip_vjp_with_plan(IcmHipPlan *plan, int Q, const double *payout, const double *v, double *equity,
                          double *dS, IcmHipRunStats *stats);
int icm_hip_vjp_with_plan_device(IcmHipPlan *plan, int Q, const double *d_payout, const double *d_v,
                                 double *d_equity, double *d_dS, IcmHipRunStats *stats);
int icm_hip_equity_vjp(int n, const double *S, int Q, const double *payout, int k, const double *v,
                       double *equity, double *dS);
/* One node: inner[i] = Psi_i and H[i] = sum_j c_j dPsi_j/da_i for given a, c
 * (players in the given order). */
int icm_hip_node_vjp(int n, const double *a, const double *c, const double *payout, int k, double *inner,
                     double *H);

/* ---- reference GPU names (src/gpu/icm_gpu.h:66-111), same signatures ---- */
int icm_gpu_init(int device_id);
void icm_gpu_shutdown(void);
const char *icm_gpu_last_error(void);
int icm_gpu_release_pooled_memory(void);
IcmHipPlan *icm_gpu_plan_create(int n, const double *S, int k, const IcmHipOptions *opts);
void icm_gpu_plan_destroy(IcmHipPlan *plan);
int icm_gpu_plan_summary(const IcmHipPlan *plan, IcmHipPlanSummary *summary);
int icm_gpu_equity_with_plan(IcmHipPlan *plan, int Q, const double *payout, double *equity,
                             IcmHipRunStats *stats);
int icm_gpu_equity(int n, const double *S, int Q, const double *payout, int k, double *equity,
                   const IcmHipOptions *opts, IcmHipRunStats *stats);
int icm_gpu_equity_subset(int n, const double *S, int Q, const double *payout, int k,
                          double *equity, const int *targets, int n_targets,
                          const IcmHipOptions *opts, IcmHipRunStats *stats);

/* ---- calibration and diagnostics (src/gpu/icm_gpu.h:112-116) ----
 * Return 1 on success, 0 on failure, as the reference.
 * hbm: bytes copied per second by 16 device-to-device copies of 1 GiB
 *   (each byte counted once, as gpu_api.cu:837-860; HBM traffic is twice it).
 * fused pair: median device time per parent of one LDS-level build launch and
 *   one LDS-level propagate launch over `batch` parents at transform size
 *   fft_n (16 .. 4096, power of two), children of fft_n/2 + 1 coefficients;
 *   the r2c variant reports the same kernels (they already pack two real
 *   sequences per complex transform). */
int icm_gpu_measure_hbm_bandwidth_gbps(double *gbps_out);
int icm_gpu_measure_fused_pair_ns(int fft_n, int batch, int quick, double *build_ns_out, double *corr_ns_out);
int icm_gpu_measure_fused_r2c_pair_ns(int fft_n, int batch, int quick, double *build_ns_out,
                                      double *corr_ns_out);
/* Streaming rates of this library's own kernels over `bytes` per array
 * (no reference counterpart): read-only (sum), write-only (fill) and copy,
 * each in HBM bytes moved per second (a copy counts its reads and writes). */
int icm_hip_measure_stream_gbps(size_t bytes, double *read_gbps, double *write_gbps, double *copy_gbps);
/* LDS read rate of conflict-free 16-byte reads (ds_read_b128) over all CUs,
 * bytes delivered to lanes per second (no reference counterpart). */
int icm_hip_measure_lds_gbps(double *gbps);

/* ---- no device allocation ----
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): computes what icm_hip_equity computes, for instances
 * that fit the one-launch small-instance kernel, with all device memory taken
 * from a static device array of the library (8.5 MB per GPU, held until the
 * process ends).  Limits: n <= 4096, Q <= 2048, and the tree must fit one
 * workgroup's LDS (about 150 KB: e.g. k <= 100 at n = 1024); larger inputs
 * return -1 with a message.  Host pointers, default stream, returns after
 * the result is copied back.  Calls that overlap on one GPU race (they share
 * that GPU's array). */
int icm_hip_equity_noalloc(int n, const double *S, int Q, const double *payout, int k, double *equity);

/* ---- CPU-named aliases bound by python/icm/__init__.py:92-119 ---- */
void icm_init(const char *wisdom_path);        /* wisdom_path ignored; selects device 0 */
void icm_equity(int n, const double *S, int Q, const double *payout, int k, double *equity);
void icm_equity_subset(int n, const double *S, int Q, const double *payout, int k,
                       double *equity, const int *targets, int n_targets);

/* ---- host-only planner query (no GPU needed) ----
 * Fills per-level geometry for (n, k, B): n_real, coefficient count c,
 * g length G, and transform size F (F[0] = 0).  Arrays need max_levels
 * entries.  Returns the level count L, or -1 on invalid input. */
int icm_hip_plan_geometry(int n, int k, int B, int max_levels, int *n_real, int *c, int *G, int *F);

/* erfc-trap nodes (src/cpu/icm.c:117-132), host-only. */
void icm_hip_make_nodes(int Q, double Smax, double *logv, double *w);

/* ---- per-kernel timing (plays the role of the reference's instrumented
 * run, IcmGpuRunStats phase fields src/gpu/icm_gpu.h:36-50, filled per kernel
 * instead of per phase) ----
 * on != 0 brackets every launch of later runs with events on the plan's
 * stream and clears the aggregates; kernel_stat reports, per kernel name,
 * launches, summed device time and summed compulsory bytes / FLOPs. */
int icm_hip_plan_set_profiling(IcmHipPlan *plan, int on);
/* Planner override (the autotuner's knob; replaces the reference's per-device
 * dispatch calibration, devices/b200/gpu_fft_config.h): levels whose children
 * keep <= cmax coefficients run as direct (schoolbook) sums instead of
 * transforms (cmax <= 512 takes effect; 0 = transforms only).  The default
 * comes from the MI355X tuning table. */
int icm_hip_plan_set_schoolbook_max(IcmHipPlan *plan, int cmax);
int icm_hip_plan_get_schoolbook_max(const IcmHipPlan *plan);
/* Small instances in one launch (the counterpart of the reference's
 * single-kernel path, src/gpu/gpu_kernels.cu:1547-1668, dispatched at
 * src/gpu/gpu_api.cu:488-610): one workgroup per quadrature node holds the
 * whole tree in LDS; used when it fits (about 150 KB) and the MI355X tuning
 * table's crossover allows.  mode -1 = tuned (default), 0 = never, 1 =
 * whenever it fits.  get returns 1 when forward runs take this path. */
int icm_hip_plan_set_small_path(IcmHipPlan *plan, int mode);
/* Column fusion of the four-step levels (default all on): a level's column
 * inverse writes the neighbouring level's column transforms directly instead
 * of coefficients that the neighbour reads back and transforms.  on: bit 1 the
 * propagate (g pairs of the level below), bit 2 the build (child pairs of the
 * level above, into the spectrum cache), bit 4 the top build level into the
 * root's child pairs, bit 8 the F = 8192 level into the level above; -1 = all.
 * Results are bit-identical either way; off is for testing and timing. */
int icm_hip_plan_set_column_fusion(IcmHipPlan *plan, int on);

/* Diagnostic, no counterpart in the reference: on != 0 fills every device
 * buffer a q-batch writes before it reads (block and level products, g
 * arrays, four-step scratch, spectrum caches, per-chunk partials; for the
 * adjoint also its tangent and B arrays) with NaN before each q-batch of
 * later runs, so that a read of a position the batch has not written itself
 * shows up as NaN in the result.  Results are otherwise unchanged; the fills
 * cost one pass over those buffers per q-batch. */
/* Adjoint memory, no counterpart in the reference (it has no adjoint):
 * interval m >= 2 keeps the tangent products of tree levels 0, m, 2m, ...
 * and recomputes each segment between two checkpoints from the lower one
 * when the reverse sweep reaches it (one extra tangent build per pass, no
 * tangent spectrum cache); 0 keeps every level; -1 (default, or the
 * ICM_HIP_ADJOINT_CHECKPOINT environment variable at plan creation) keeps
 * every level when that leaves at least min(q-batch, 32) q-planes per
 * adjoint batch and otherwise takes the interval that leaves the most.
 * Results are bit-identical for every interval.  Changing it frees the
 * adjoint buffers; the next adjoint call sizes them again. */
int icm_hip_plan_set_adjoint_checkpoint(IcmHipPlan *plan, int interval);
/* Interval in use (0: every level kept) and q-planes per adjoint batch,
 * after the first adjoint call; -1 before it. */
int icm_hip_plan_adjoint_layout(const IcmHipPlan *plan, int *interval, int *q_planes);

int icm_hip_plan_set_poison(IcmHipPlan *plan, int on);
/* Bytes NaN-filled by the plan's last run (0 with poisoning off; -1 for a
 * null plan). */
long long icm_hip_plan_poisoned_bytes(const IcmHipPlan *plan);

/* Diagnostic, no counterpart in the reference: checks on the host (no
 * device needed) the index maps the four-step row-pair kernels rely on --
 * every LDS layout one to one, the row pairs of the workgroups covering every
 * row once, and each workgroup's representatives covering both tiles once
 * with mirror-bin pairs.  0 when they hold, -1 with the first violation in
 * icm_hip_last_error. */
int icm_hip_selfcheck_row_layouts(void);
int icm_hip_plan_get_small_path(const IcmHipPlan *plan);
/* Spectrum cache (no counterpart in the reference, whose large levels are
 * cuFFT plans per level, src/gpu/gpu_api.cu): the column transforms of each
 * four-step build level's child pairs stay in HBM until the propagate of the
 * same q-batch, which then does not transform the children again.  Planned on
 * by default when at least min(q-batch, 32) q-planes fit with it; on = 0
 * frees it, on = 1 allocates it again (-1 when it does not fit).  Pruned
 * subset runs read it through their parent lists. */
int icm_hip_plan_set_spectrum_cache(IcmHipPlan *plan, int on);
int icm_hip_plan_get_spectrum_cache(const IcmHipPlan *plan);
/* Every profiled launch in order: kernel name, tree level (l > 0 build of
 * level l, -l propagate of level l, 0 leaves and reductions), device ns and
 * compulsory bytes. */
int icm_hip_plan_launch_count(const IcmHipPlan *plan);
int icm_hip_plan_launch_stat(const IcmHipPlan *plan, int index, char *name, int name_len, int *level, double *ns,
                             double *alg_bytes);
int icm_hip_plan_kernel_count(const IcmHipPlan *plan);
int icm_hip_plan_kernel_stat(const IcmHipPlan *plan, int index, char *name, int name_len, int *launches,
                             double *total_ns, double *alg_bytes, double *alg_flops);

/* ---- ragged batches (the batch entry with offset arrays that this engine
 * adds beside the reference's single-instance API; no reference counterpart)
 * n_inst independent instances.  Instance i owns players n_off[i] ..
 * n_off[i+1]-1 of S and payout entries k_off[i] .. k_off[i+1]-1 of payout
 * (n_off[0] = k_off[0] = 0, 1 <= k_i <= n_i); equities come back in the flat
 * order of S.  Each instance gets what icm_equity(n_i, S_i, Q, payout_i,
 * k_i, equity_i) computes, with its own nodes (its own max stack).
 * Instances whose tree levels all fit one LDS transform (<= 4096 points:
 * k_i <= 2048, or n_i <= 4096) share one launch per tree level; the others
 * (four-step levels) each run through a single-instance plan inside the
 * batch, planned for an equal share of the memory fraction.
 * opts: NULL or enable_graphs != 0
 * records the first run for each Q as one HIP graph and replays it;
 * verbose; the other fields as for single plans.  Calls on one batch plan
 * must not overlap. */
typedef struct IcmHipBatchPlan IcmHipBatchPlan;
IcmHipBatchPlan *icm_hip_batch_plan_create(int n_inst, const int *n_off, const double *S, const int *k_off,
                                           const IcmHipOptions *opts);
/* With IcmHipPlanParams (block size, q-planes per pass, memory budget). */
IcmHipBatchPlan *icm_hip_batch_plan_create_ex(int n_inst, const int *n_off, const double *S, const int *k_off,
                                              const IcmHipOptions *opts, const IcmHipPlanParams *params);
void icm_hip_batch_plan_destroy(IcmHipBatchPlan *plan);
int icm_hip_batch_equity(IcmHipBatchPlan *plan, int Q, const double *payout, double *equity,
                         IcmHipRunStats *stats);
/* d_payout[K] and d_equity[N] are device pointers (K = k_off[n_inst], N = n_off[n_inst]). */
int icm_hip_batch_equity_device(IcmHipBatchPlan *plan, int Q, const double *d_payout, double *d_equity,
                                IcmHipRunStats *stats);
/* Adjoint of a ragged batch: equities and dS = d(v.E)/dS per instance for a
 * cotangent v (flat, same order as S); each instance as icm_hip_vjp_with_plan.
 * equity / dS may be NULL.  The first call allocates the adjoint buffers;
 * the first call per Q is recorded as its own HIP graph. */
int icm_hip_batch_vjp(IcmHipBatchPlan *plan, int Q, const double *payout, const double *v, double *equity,
                      double *dS, IcmHipRunStats *stats);
int icm_hip_batch_vjp_device(IcmHipBatchPlan *plan, int Q, const double *d_payout, const double *d_v,
                             double *d_equity, double *d_dS, IcmHipRunStats *stats);
/* One-shot: plan, run (no graph), free. */
int icm_hip_equity_batch(int n_inst, const int *n_off, const double *S, const int *k_off, int Q,
                         const double *payout, double *equity);
int icm_hip_batch_summary(const IcmHipBatchPlan *plan, int *n_inst, int *n_players, int *n_blocks,
                          int *max_levels, int *q_batch, size_t *device_bytes);
/* Schoolbook / transform crossover of the batch's forward levels (as
 * icm_hip_plan_set_schoolbook_max; the adjoint keeps transforms). */
int icm_hip_batch_set_schoolbook_max(IcmHipBatchPlan *plan, int cmax);
/* Forward batch as ONE persistent launch per q-pass (k_batch_persist) instead
 * of one launch per phase and tree level: workgroups take leaf, level,
 * extract and reduction work items by ticket, and each instance starts its
 * next level as soon as its own previous level is done (per-instance
 * completion counters, no barrier between levels).  Same equities as the
 * per-level launches.  Needs block size 64 (the tuned default; -1 otherwise);
 * the default comes from ICM_HIP_BATCH_PERSISTENT (0/1).  The captured graph
 * then holds that one kernel per pass.  The adjoint keeps per-level launches.
 * set_spin_limit bounds each item's wait for its inputs (default 2^20 polls,
 * under a second): past it the run fails with an error instead of waiting on. */
int icm_hip_batch_set_persistent(IcmHipBatchPlan *plan, int on);
int icm_hip_batch_is_persistent(const IcmHipBatchPlan *plan);
int icm_hip_batch_set_spin_limit(IcmHipBatchPlan *plan, int spins);
/* Per-kernel timing of batch runs (runs eagerly, without the graph, while on). */
int icm_hip_batch_set_profiling(IcmHipBatchPlan *plan, int on);
int icm_hip_batch_kernel_count(const IcmHipBatchPlan *plan);
int icm_hip_batch_kernel_stat(const IcmHipBatchPlan *plan, int index, char *name, int name_len, int *launches,
                              double *total_ns, double *alg_bytes, double *alg_flops);

#ifdef __cplusplus
}
#endif
#endif /* ICM_HIP_H */
