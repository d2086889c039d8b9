/*
 * icm_oracle.h -- CPU restatement of the reference ICM equity algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker / CPU baseline.  The product library (libicm_hip.so) never links
 * or calls it.
 *
 * Every function cites the reference file:line it restates
 * (paths relative to the reference repository root).
 */
#ifndef ICM_ORACLE_H
#define ICM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Engines for orc_equity / orc_inner. */
#define ORC_ENGINE_NAIVE  0   /* src/cpu/icm.c:1880-1914 */
#define ORC_ENGINE_LINEAR 1   /* src/cpu/icm.c:1981-2076 (flat variant) */
#define ORC_ENGINE_HYBRID 2   /* src/cpu/icm.c:2126-2318 + tree 1205-1819 */

/* erfc-trap quadrature nodes (src/cpu/icm.c:117-132, Q==1 guard as in
 * src/gpu/gpu_exec.cu:17). */
void orc_make_nodes(int Q, double Smax, double *logv, double *w);

/* Closed forms (src/cpu/icm.c:145-170). */
void orc_v1_exact(int n, const double *S, double *V1);
void orc_v2_exact(int n, const double *S, double *V2);

/* One quadrature node: inner[i] = Psi_i = sum_{r<k} payout[r] [x^r]
 * prod_{j!=i} (a_j + (1-a_j) x), players in the given order (no sort).
 * B: hybrid block size (ignored by other engines).  fft_min: products
 * whose shorter operand has at least this many coefficients use the FFT
 * (0 = schoolbook only).  active may be NULL (all players). */
int orc_inner(int n, const double *a, const double *payout, int k,
              double *inner, int engine, int B, int fft_min,
              const unsigned char *active);

/* Full equity, following the integration loop of src/cpu/icm.c:2523-2611
 * with the hybrid engine's stack-descending sort (src/cpu/icm.c:2119-2143,
 * index tie-break as src/gpu/gpu_api.cu:492-502).
 *   targets==NULL: all players; else subset semantics of
 *   src/cpu/icm.c:2475-2521 (non-targets are 0, cold subtrees pruned).
 *   q_lo/q_hi: only nodes q in [q_lo, q_hi) are integrated (a bounded
 *   sample for CPU timing); pass 0, Q for the whole rule.
 *   nthreads: OpenMP threads over quadrature nodes (<=0: default).
 * Returns 0 on success, -1 on invalid arguments. */
int orc_equity(int n, const double *S, int Q, const double *payout, int k,
               double *equity, const int *targets, int n_targets,
               int engine, int B, int fft_min, int q_lo, int q_hi,
               int nthreads);

#ifdef __cplusplus
}
#endif
#endif
