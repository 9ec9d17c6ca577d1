/*
 * mfgpu.h -- C ABI of the MI355X numeric phase for PureUMFPACK.jl's supernodal
 * multifrontal LU.
 *
 * The reference has no FFI: its boundary is the Julia call surface
 *   splu(A; method=:multifrontal, scale, check)      src/interface.jl:93-120
 *   multifrontal_lu(A; q, tol, check)                src/multifrontal.jl:146-149
 *   solve(F::PureLU, b; refine), F \ b               src/interface.jl:170-179
 * Julia keeps ordering, scaling and symbolic analysis (src/amd.jl, src/scaling.jl,
 * src/symbolic.jl) and hands their results to mf_plan_create; every numeric step of
 * the supernode loop (src/multifrontal.jl:198-357), the triangular solves
 * (src/solve.jl:10-39, src/interface.jl:123-138) and the refinement loop
 * (src/interface.jl:170-177) runs on one GPU behind the calls below.
 *
 * Conventions
 *   - index arrays are 1-based int64 (Julia's Vector{Int}); values are column-major;
 *   - the caller owns every host array; the library never keeps a host pointer after a
 *     call returns; device memory belongs to the plan and is freed by mf_plan_destroy;
 *   - functions return MF_OK (0) or one of the error codes; mf_last_error() gives the
 *     text of the calling thread's last error;
 *   - a plan is not re-entrant: calls on one plan must be serialised by the caller.
 *     Calls are synchronous from the caller's point of view.
 */
#ifndef MFGPU_H
#define MFGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum mf_status {
    MF_OK = 0,
    MF_SINGULAR = 1, /* exactly-zero block pivot with check != 0: *info_col is the column */
    MF_DIM = 2,      /* size mismatch (DimensionMismatch in the reference) */
    MF_BADARG = 3,   /* bad option / argument (ArgumentError) */
    MF_HIP = 4,      /* HIP runtime error */
    MF_OOM = 5       /* device allocation failed */
};

enum mf_precision {
    MF_FP64 = 0, /* parity path: fronts, getrf, trsm, Schur gemm in FP64 */
    MF_FP32 = 1, /* fronts in FP32 (FP32 accumulate); FP64 residual and update */
    MF_BF16 = 2  /* FP32 fronts/getrf/trsm, Schur gemm operands in BF16 with per-tile
                    power-of-two scaling, FP32 accumulate; FP64 residual and update */
};

enum mf_ordering { MF_ORDER_AMD = 0, MF_ORDER_COLAMD = 1, MF_ORDER_GIVEN = 2 };

enum mf_pivoting {
    MF_PIVOT_AUTO = 0,    /* large fronts: interchange-free blocks with a multiplier check,
                             redone with partial pivoting if the check fails */
    MF_PIVOT_PARTIAL = 1  /* large fronts: always the partial-pivoting panel schedule */
};

enum mf_solve_path {
    MF_SOLVE_AUTO = 0,    /* column-parallel sweeps from 512 right-hand-side columns (matrices
                             times right-hand sides) when no pivot block exceeds 1024 columns,
                             the per-front sweeps otherwise */
    MF_SOLVE_FRONTS = 1,  /* per-front sweeps: a front's rows spread over the lanes */
    MF_SOLVE_COLUMNS = 2  /* column-parallel sweeps: one lane per right-hand-side column */
};

typedef struct mf_options {
    int32_t precision; /* enum mf_precision */
    int32_t small_nf;  /* fronts with nf <= small_nf take the one-workgroup fused kernel;
                          0 = library default */
    int32_t use_graph; /* 1: capture the factor schedule and the solve sweeps in hipGraphs and
                          replay them (same kernels in the same order: identical results) */
    int32_t nbatch;    /* matrices sharing the pattern that one plan holds (>= 1) */
    int32_t pivoting;  /* enum mf_pivoting */
    float pivot_tol;   /* MF_PIVOT_AUTO accepts the diagonal while every multiplier satisfies
                          |l_ij| <= 1/pivot_tol.  0 = default: 1 for MF_FP64 (then the result
                          is exactly that of partial pivoting, src/multifrontal.jl:248), 0.1 for
                          the mixed precisions (threshold pivoting) */
    int32_t solve_path; /* enum mf_solve_path: how mf_solve* sweep the factors */
    int32_t reserved;
} mf_options;

typedef struct mf_solve_report {
    int32_t iterations;     /* refinement steps taken */
    int32_t converged;      /* 1 if bwd_tol > 0 and the bound was met */
    double backward_error;  /* max over RHS of ||b-Ax||_inf / (||A||_inf ||x||_inf + ||b||_inf) */
    double residual_2;      /* max over RHS of ||b-Ax||_2 / ||b||_2 (src test metric) */
    double seconds;         /* device time of the whole call */
} mf_solve_report;

/* ---- host symbolic analysis (the reference keeps this in Julia; offered for non-Julia
 *      hosts and for tests).  Restates symbolic_mf (src/symbolic.jl:211-226) with AMD on
 *      A+A^T (src/amd.jl:141-148) or on A^T A (colamd_order, src/amd.jl:183-188). ---- */
typedef struct mf_symbolic mf_symbolic;

/* Ap/Ai: CSC pattern of A (1-based).  ordering = MF_ORDER_GIVEN uses q (1-based, length n)
 * as the pre-postorder column order, like symbolic_mf(A; q). */
int mf_analyze(int64_t n, const int64_t *Ap, const int64_t *Ai, int32_t ordering,
               const int64_t *q, mf_symbolic **out);
/* Sizes: nsuper, total update-row entries (sum of |colstruct[lastcol(s)]|), total child
 * entries, and sum over all columns of |colstruct[j]| (predicted_fill = 2(n + ncolstruct)). */
int mf_symbolic_sizes(const mf_symbolic *S, int64_t *nsuper, int64_t *nupd, int64_t *nchild,
                      int64_t *ncolstruct);
/* Copies the analysis out (all 1-based): qf (n), parent (n, 0 = root), sstart (nsuper+1),
 * upd_ptr (nsuper+1) / upd_idx (nupd): colstruct of each supernode's last column,
 * child_ptr (nsuper+1) / child_idx (nchild): childsupers, ascending.  Any pointer may be
 * NULL to skip that array. */
int mf_symbolic_export(const mf_symbolic *S, int64_t *qf, int64_t *parent, int64_t *sstart,
                       int64_t *upd_ptr, int64_t *upd_idx, int64_t *child_ptr,
                       int64_t *child_idx);
/* Relaxed supernode amalgamation of an analysis, in place (opt-in; no reference counterpart:
 * the reference names it as further work, README.md:186-188, results/PERF_DIAGNOSIS.md:76-78).
 * A supernode merges into its parent when it is the child whose columns end right before the
 * parent's and either the merged pivot block has at most max_small_np columns or the explicit
 * zeros the merge stores are at most max_zero_fraction of the merged supernode's L and U
 * entries.  Fewer, larger fronts and a shallower tree; L and U then hold explicit zeros, so
 * the pattern no longer equals the reference's (mf_symbolic_export gives the relaxed arrays and
 * sizes).  *merged = number of merges. */
int mf_symbolic_relax(mf_symbolic *S, double max_zero_fraction, int64_t max_small_np, int64_t *merged);
void mf_symbolic_destroy(mf_symbolic *S);

/* ---- the numeric phase ---- */
typedef struct mf_plan mf_plan;

/* Builds the static device schedule from SymbolicMF-shaped arrays (src/symbolic.jl:203-209)
 * and A's pattern.  Replaces the prologue of multifrontal_lu (src/multifrontal.jl:150-197)
 * and _factor_colptrs (src/multifrontal.jl:73-104). */
int mf_plan_create(int64_t n, const int64_t *qf, int64_t nsuper, const int64_t *sstart,
                   const int64_t *upd_ptr, const int64_t *upd_idx, const int64_t *child_ptr,
                   const int64_t *child_idx, const int64_t *Ap, const int64_t *Ai,
                   const mf_options *opt, mf_plan **out);

/* Numeric factorization of one matrix with A's pattern: values Ax (nnz, CSC order of the
 * pattern given to mf_plan_create), row scaling Rs (n; NULL = no scaling).  Replaces the
 * supernode loop src/multifrontal.jl:198-357.  With check != 0 an exactly-zero block
 * pivot returns MF_SINGULAR and *info_col = its column (1-based, in factor order, as
 * SingularException(c1+info-1) at src/multifrontal.jl:250-251).  The pivot-block inverses of
 * the top level's fronts, which only the solves use, are computed by the first solve after
 * the factorization (beside its sweeps of the levels below). */
int mf_factor(mf_plan *P, const double *Ax, const double *Rs, int32_t check,
              int64_t *info_col);
/* Same with Ax and Rs already in device memory (HBM-resident inputs). */
int mf_factor_device(mf_plan *P, const double *d_Ax, const double *d_Rs, int32_t check,
                     int64_t *info_col);

/* Same-pattern batch: values of nmat matrices (nmat x nnz, row-major by matrix) and their
 * row scalings (nmat x n, or NULL).  nmat <= opt.nbatch.  info_col[m] = first singular
 * column of matrix m (1-based) or 0.  No reference counterpart (the reference refactors
 * one matrix per call, src/multifrontal.jl:146). */
int mf_refactor_batch(mf_plan *P, int64_t nmat, const double *Ax_batch, const double *Rs_batch,
                      int64_t *info_col);
/* Same with the values and scalings in device memory. */
int mf_refactor_batch_device(mf_plan *P, int64_t nmat, const double *d_Ax_batch, const double *d_Rs_batch,
                             int64_t *info_col);

/* Solve A X = B for nrhs right-hand sides (B: n x nrhs, leading dimension ldb; X: ldx).
 * refine_steps >= 0 fixed refinement steps as solve(F, b; refine=k)
 * (src/interface.jl:170-177); with bwd_tol > 0 the loop stops early once the normwise
 * backward error is <= bwd_tol (refine_steps is then the maximum).  rep may be NULL. */
int mf_solve(mf_plan *P, int64_t nrhs, const double *B, int64_t ldb, double *X, int64_t ldx,
             int32_t refine_steps, double bwd_tol, mf_solve_report *rep);
/* Same with B and X in device memory. */
int mf_solve_device(mf_plan *P, int64_t nrhs, const double *d_B, int64_t ldb, double *d_X,
                    int64_t ldx, int32_t refine_steps, double bwd_tol, mf_solve_report *rep);
/* mf_solve_device that allocates nothing: no reference counterpart.  It is for callers that
 * cannot allocate device memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The solve's vectors live in a file-scope device array of 64 MiB (one copy
 * per GPU, held until the process ends); a solve needing more (about 7 n nrhs doubles plus
 * the update vectors) returns MF_BADARG with the size in mf_last_error().  Calls on one GPU
 * share the array, so calls that overlap on one GPU (concurrent plans) race.  Profiling must be
 * off.  Same arguments and results as mf_solve_device. */
int mf_solve_device_noalloc(mf_plan *P, int64_t nrhs, const double *d_B, int64_t ldb, double *d_X,
                            int64_t ldx, int32_t refine_steps, double bwd_tol, mf_solve_report *rep);
/* Batched solve after mf_refactor_batch: matrix m solves B[m] (n x nrhs, contiguous). */
int mf_solve_batch(mf_plan *P, int64_t nmat, int64_t nrhs, const double *B, double *X,
                   int32_t refine_steps, double bwd_tol, mf_solve_report *rep);
/* Same with B and X in device memory. */
int mf_solve_batch_device(mf_plan *P, int64_t nmat, int64_t nrhs, const double *d_B, double *d_X,
                          int32_t refine_steps, double bwd_tol, mf_solve_report *rep);

/* nnz(L) and nnz(U) of the factor (each = predicted_fill/2, src/symbolic.jl:230-237). */
int mf_nnz(const mf_plan *P, int64_t *nnzL, int64_t *nnzU);

/* Exports L, U (CSC, 1-based, exactly the reference's arrays: src/multifrontal.jl:277-350),
 * p = qf[prow] and q = qf (src/multifrontal.jl:355).  Lnzval/Unzval are FP64 (converted
 * from the front precision).  Any pointer may be NULL. */
int mf_export_factors(mf_plan *P, int64_t *Lcolptr, int64_t *Lrowval, double *Lnzval,
                      int64_t *Ucolptr, int64_t *Urowval, double *Unzval, int64_t *p,
                      int64_t *q);
/* mf_export_factors for matrix mat (0-based) of the last mf_refactor_batch.  No reference
 * counterpart (batches are an extension of src/multifrontal.jl:146). */
int mf_export_factors_batch(mf_plan *P, int64_t mat, int64_t *Lcolptr, int64_t *Lrowval, double *Lnzval,
                            int64_t *Ucolptr, int64_t *Urowval, double *Unzval, int64_t *p, int64_t *q);

/* Statistics of the plan: factor flops, Schur-gemm flops, flops in fronts with nf >= 1024,
 * device bytes of the fronts, number of levels, kernel launches per factorization. */
typedef struct mf_plan_stats {
    int64_t last_schedule;   /* 0: some matrix of the last call ran the pivoted panels,
                                1: the interchange-free blocks held for every matrix */
    double factor_flops;
    double gemm_flops;
    double gemm_flops_big;   /* Schur gemm flops of fronts with nf >= 1024 */
    double front_bytes;      /* store_bytes + arena_bytes */
    int64_t nsuper;
    int64_t nlevels;
    int64_t launches;
    int64_t max_nf;
    double store_bytes;      /* factor store: L, U12 and the pivot-block inverses (all matrices) */
    double arena_bytes;      /* workspace arena: fronts and contribution blocks, liveness-planned */
    double resident_front_bytes; /* what every nf x nf front kept resident would need */
    int64_t last_redone;     /* matrices the last call factored on the pivoted panels */
} mf_plan_stats;
int mf_plan_get_stats(const mf_plan *P, mf_plan_stats *st);
/* The statistics mf_plan_create would give (sizes of the factor store and the workspace
 * arena, flops, launches), computed on the host without touching the GPU: for sizing a
 * problem before committing device memory.  last_schedule reports the schedule a call would
 * try first.  No reference counterpart. */
int mf_plan_estimate(int64_t n, const int64_t *qf, int64_t nsuper, const int64_t *sstart,
                     const int64_t *upd_ptr, const int64_t *upd_idx, const int64_t *child_ptr,
                     const int64_t *child_idx, const int64_t *Ap, const int64_t *Ai,
                     const mf_options *opt, mf_plan_stats *st);

/* Per-kernel-class timing with HIP events on the plan's stream (instrumentation; no
 * reference counterpart).  While profiling is on, mf_factor* and mf_solve* record an event
 * pair around every launch and add the elapsed time to its class:
 *   0 fused small fronts, 1 column assembly, 2 panel getrf, 3 panel-update gemm,
 *   4 Schur gemm, 5 forward solve, 6 backward solve, 7 residual/norms/vector kernels.
 * flops[k] / bytes[k] are the algorithmic flops and HBM bytes of class k accumulated over the
 * same launches.  Profiling disables graph replay. */
#define MF_PROF_CLASSES 8
int mf_plan_set_profiling(mf_plan *P, int32_t on);
int mf_plan_get_profile(const mf_plan *P, double *ms, int64_t *launches, double *flops, double *bytes);

const char *mf_last_error(void);
void mf_plan_destroy(mf_plan *P);

/* ---- gemm tile policy and the autotuner's measurement (no reference counterpart: tile
 * shapes are a property of the GPU kernels; every shape gives bit-identical results) ---- */
/* policy[use * nbuckets + bucket] = tile index for use 0 (panel updates) and 1 (Schur
 * complements), bucket = size class of sqrt(m n); applies to plans created afterwards. */
int mf_set_gemm_tiles(int32_t precision, const int32_t *policy);
int mf_get_gemm_tiles(int32_t precision, int32_t *policy);
/* tile shapes (rows x columns per workgroup) and bucket upper bounds (nbuckets - 1 values). */
int mf_gemm_tile_info(int32_t *ntiles, int32_t *tile_m, int32_t *tile_n, int32_t *nbuckets, int32_t *bucket_hi);
/* average seconds of one launch of nprob independent C(m x n) -= A(m x k) B(k x n) problems
 * with tile shape `tile` (reps launches timed with HIP events after one warm-up). */
int mf_time_gemm(int32_t precision, int64_t m, int64_t n, int64_t k, int64_t nprob, int32_t tile, int32_t reps,
                 double *seconds);

/* average seconds of one launch of the nprob problems mnk[3q..3q+2] = (m, n, k) (independent
 * C -= A B, as a plan groups a level's Schur complements) with tile shape `tile`. */
int mf_time_gemm_list(int32_t precision, int64_t nprob, const int32_t *mnk, int32_t tile, int32_t reps,
                      double *seconds);

/* Copies `bytes` (a multiple of 16) between two fresh device buffers with a plain 16-byte
 * copy kernel: the known traffic for calibrating hardware byte counters. */
int mf_copy_probe(int64_t bytes, double *seconds);

/* ---- small MLP cost model, trained on the GPU (no reference counterpart) ----
 * d_in -> hidden (ReLU) -> hidden (ReLU) -> 1, FP32; samples row-major (n x d_in); loss =
 * mean squared error; parameters: per layer W (d_out x d_in, row-major) then b. */
typedef struct mf_mlp mf_mlp;
int mf_mlp_create(int32_t d_in, int32_t hidden, uint64_t seed, mf_mlp **out);
void mf_mlp_destroy(mf_mlp *m);
int64_t mf_mlp_num_params(const mf_mlp *m);
int mf_mlp_get_params(mf_mlp *m, float *params);
int mf_mlp_set_params(mf_mlp *m, const float *params);
int mf_mlp_forward(mf_mlp *m, const float *X, int64_t n, float *Y);
/* loss and its gradient with respect to every parameter (same layout as the parameters) */
int mf_mlp_gradient(mf_mlp *m, const float *X, const float *t, int64_t n, float *loss, float *grad);
/* full-batch Adam (lr, beta1 0.9, beta2 0.999, eps 1e-8) for `epochs` steps */
int mf_mlp_train(mf_mlp *m, const float *X, const float *t, int64_t n, int32_t epochs, float lr, float *final_loss);
const char *mf_mlp_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
