/*
 * mf_oracle.c -- CPU restatement of the numeric multifrontal LU path of
 * PureUMFPACK.jl.  TEST INFRASTRUCTURE ONLY: built and loaded by oracle/numeric.py,
 * used by tests/, by the smoke check and by the CPU-baseline leg of bench.py.  The
 * product library never links or calls it.
 *
 * What it restates (0-based indices throughout; the reference is 1-based):
 *   mfo_colptrs  L/U column counts                       src/multifrontal.jl:73-104
 *   mfo_factor   the supernode loop: front assembly of V and V^T entries (:216-231),
 *                extend-add of children in reverse order (:45-57, :235-243),
 *                getrf of the np x np pivot block with row-maximum partial pivoting
 *                restricted to the block (:247-252), row swaps of R12 (:256-263),
 *                L21 <- A21 U11^-1 and U12 <- L11^-1 R12 (:264-266), the Schur update
 *                F22 <- F22 - L21 U12 (:267-268), pivot rows (:272-275, :363-375), the
 *                direct CSC scatter of L (V-row ids) and U (:277-317), the CB push
 *                (:319-332) and the final relabel of L rows through invperm(prow) (:347-350).
 *   mfo_lsolve   src/solve.jl:10-22;  mfo_usolve src/solve.jl:25-39
 *   mfo_solve_factor src/interface.jl:123-138 (Rs may be NULL: the raw-factor solve of
 *                src/solve.jl:66-77)
 *   mfo_spmv_residual r = b - A x with a column-oriented CSC SpMV (src/interface.jl:173)
 *
 * Dense kernels: backend 0 runs the reference-BLAS loop orders (dgetf2 with idamax
 * first-maximum tie break and reciprocal scaling, dtrsm R/U/N/N and L/L/N/U, dgemm N/N);
 * backend 1 calls LAPACK/BLAS dgetrf/dtrsm/dgemm through function pointers handed in by
 * mfo_set_blas (numpy's bundled OpenBLAS, ILP64), the same calls the reference makes
 * through Julia's LinearAlgebra.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t i64;

typedef void (*dgetrf_fn)(const i64 *, const i64 *, double *, const i64 *, i64 *, i64 *);
typedef void (*dtrsm_fn)(const char *, const char *, const char *, const char *, const i64 *,
                         const i64 *, const double *, const double *, const i64 *, double *,
                         const i64 *, size_t, size_t, size_t, size_t);
typedef void (*dgemm_fn)(const char *, const char *, const i64 *, const i64 *, const i64 *,
                         const double *, const double *, const i64 *, const double *, const i64 *,
                         const double *, double *, const i64 *, size_t, size_t);

static dgetrf_fn p_getrf = 0;
static dtrsm_fn p_trsm = 0;
static dgemm_fn p_gemm = 0;

void mfo_set_blas(void *getrf, void *trsm, void *gemm)
{
    p_getrf = (dgetrf_fn)getrf;
    p_trsm = (dtrsm_fn)trsm;
    p_gemm = (dgemm_fn)gemm;
}

/* ---------------- reference-BLAS loop orders (backend 0) ---------------- */

/* dgetf2: ipiv 0-based, returns LAPACK info (1-based first exactly-zero pivot, 0 if none). */
static i64 getf2_loops(i64 m, double *A, i64 lda, i64 *ipiv)
{
    i64 info = 0;
    for (i64 j = 0; j < m; j++) {
        double *cj = A + j * lda;
        i64 p = j;
        double amax = fabs(cj[j]);
        for (i64 i = j + 1; i < m; i++) {
            double v = fabs(cj[i]);
            if (v > amax) { amax = v; p = i; }
        }
        ipiv[j] = p;
        if (cj[p] != 0.0) {
            if (p != j)
                for (i64 c = 0; c < m; c++) {
                    double t = A[j + c * lda]; A[j + c * lda] = A[p + c * lda]; A[p + c * lda] = t;
                }
            if (fabs(cj[j]) >= DBL_MIN) {
                double r = 1.0 / cj[j];
                for (i64 i = j + 1; i < m; i++) cj[i] *= r;
            } else {
                for (i64 i = j + 1; i < m; i++) cj[i] /= cj[j];
            }
        } else if (info == 0) {
            info = j + 1;
        }
        for (i64 c = j + 1; c < m; c++) {
            double u = A[j + c * lda];
            if (u != 0.0) {
                double *cc = A + c * lda;
                for (i64 i = j + 1; i < m; i++) cc[i] -= cj[i] * u;
            }
        }
    }
    return info;
}

/* B (m x n) <- B * inv(U), U upper non-unit n x n (dtrsm R,U,N,N). */
static void trsm_run_loops(i64 m, i64 n, const double *U, i64 ldu, double *B, i64 ldb)
{
    for (i64 j = 0; j < n; j++) {
        double *bj = B + j * ldb;
        for (i64 k = 0; k < j; k++) {
            double u = U[k + j * ldu];
            if (u != 0.0) {
                const double *bk = B + k * ldb;
                for (i64 i = 0; i < m; i++) bj[i] -= u * bk[i];
            }
        }
        double t = 1.0 / U[j + j * ldu];
        for (i64 i = 0; i < m; i++) bj[i] = t * bj[i];
    }
}

/* B (m x n) <- inv(L) * B, L unit lower m x m (dtrsm L,L,N,U). */
static void trsm_llu_loops(i64 m, i64 n, const double *L, i64 ldl, double *B, i64 ldb)
{
    for (i64 j = 0; j < n; j++) {
        double *bj = B + j * ldb;
        for (i64 k = 0; k < m; k++) {
            double bk = bj[k];
            if (bk != 0.0) {
                const double *lk = L + k * ldl;
                for (i64 i = k + 1; i < m; i++) bj[i] -= bk * lk[i];
            }
        }
    }
}

/* C (m x n) <- C - A (m x k) * B (k x n)  (dgemm N,N, alpha -1, beta 1). */
static void gemm_loops(i64 m, i64 n, i64 k, const double *A, i64 lda, const double *B, i64 ldb,
                       double *C, i64 ldc)
{
    for (i64 j = 0; j < n; j++) {
        double *cj = C + j * ldc;
        for (i64 l = 0; l < k; l++) {
            double t = -B[l + j * ldb];
            if (t != 0.0) {
                const double *al = A + l * lda;
                for (i64 i = 0; i < m; i++) cj[i] += t * al[i];
            }
        }
    }
}

/* ---------------- column pointers (src/multifrontal.jl:73-104) ---------------- */

void mfo_colptrs(i64 n, i64 nsuper, const i64 *sstart, const i64 *upd_ptr, const i64 *upd_idx,
                 i64 *Lp, i64 *Up)
{
    i64 *ucount = (i64 *)calloc((size_t)(n > 0 ? n : 1), sizeof(i64));
    for (i64 s = 0; s < nsuper; s++) {
        i64 c1 = sstart[s], c2 = sstart[s + 1] - 1, np = c2 - c1 + 1;
        i64 nu = upd_ptr[s + 1] - upd_ptr[s];
        for (i64 j = c1; j <= c2; j++) {
            Lp[j + 1] = 1 + (c2 - j) + nu;
            ucount[j] += j - c1 + 1;
        }
        for (i64 t = 0; t < nu; t++) ucount[upd_idx[upd_ptr[s] + t]] += np;
    }
    Lp[0] = 0;
    for (i64 j = 0; j < n; j++) Lp[j + 1] += Lp[j];
    Up[0] = 0;
    for (i64 j = 0; j < n; j++) Up[j + 1] = Up[j] + ucount[j];
    free(ucount);
}

/* ---------------- numeric factorization ---------------- */

/*
 * V (n x n CSC, rows sorted) is As[qf, qf].  Supernode s owns columns sstart[s]..sstart[s+1]-1,
 * its update rows are upd_idx[upd_ptr[s] .. upd_ptr[s+1]) (= colstruct of its last column,
 * ascending), its children child_idx[child_ptr[s] .. child_ptr[s+1]) (ascending).
 * Outputs: Lp/Li/Lx, Up/Ui/Ux (CSC, 0-based, L rows already in factor order), prow
 * (factor row -> V row).  Returns 0, or 1 with *info_col = first singular column
 * (0-based) when check != 0 and a block pivot is exactly zero; with check == 0 the
 * factorization runs to the end and *info_col holds the first such column or -1.
 */
int mfo_factor(i64 n, const i64 *Vp, const i64 *Vi, const double *Vx, i64 nsuper,
               const i64 *sstart, const i64 *upd_ptr, const i64 *upd_idx, const i64 *child_ptr,
               const i64 *child_idx, int backend, int check, i64 *Lp, i64 *Li, double *Lx,
               i64 *Up, i64 *Ui, double *Ux, i64 *prow, i64 *info_col)
{
    *info_col = -1;
    /* V^T (row-wise access for the strict upper part, src/multifrontal.jl:155) */
    i64 nnz = Vp[n];
    i64 *Tp = (i64 *)calloc((size_t)n + 1, sizeof(i64));
    i64 *Ti = (i64 *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(i64));
    double *Tx = (double *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
    for (i64 p = 0; p < nnz; p++) Tp[Vi[p] + 1]++;
    for (i64 i = 0; i < n; i++) Tp[i + 1] += Tp[i];
    {
        i64 *cur = (i64 *)malloc((size_t)(n > 0 ? n : 1) * sizeof(i64));
        memcpy(cur, Tp, (size_t)n * sizeof(i64));
        for (i64 j = 0; j < n; j++)
            for (i64 p = Vp[j]; p < Vp[j + 1]; p++) {
                i64 q = cur[Vi[p]]++;
                Ti[q] = j;
                Tx[q] = Vx[p];
            }
        free(cur);
    }

    i64 maxnf = 0;
    for (i64 s = 0; s < nsuper; s++) {
        i64 nf = (sstart[s + 1] - sstart[s]) + (upd_ptr[s + 1] - upd_ptr[s]);
        if (nf > maxnf) maxnf = nf;
    }
    i64 *loc = (i64 *)malloc((size_t)(n > 0 ? n : 1) * sizeof(i64));
    for (i64 i = 0; i < n; i++) { loc[i] = -1; prow[i] = i; }
    double *Fbuf = (double *)malloc((size_t)(maxnf > 0 ? maxnf * maxnf : 1) * sizeof(double));
    i64 cbcap = maxnf * maxnf > 64 ? maxnf * maxnf : 64;
    double *cb = (double *)malloc((size_t)cbcap * sizeof(double));
    i64 cbtop = 0;
    i64 *cboff = (i64 *)malloc((size_t)(nsuper > 0 ? nsuper : 1) * sizeof(i64));
    i64 *ipiv = (i64 *)malloc((size_t)(maxnf > 0 ? maxnf : 1) * sizeof(i64));
    i64 *orig = (i64 *)malloc((size_t)(maxnf > 0 ? maxnf : 1) * sizeof(i64));

    mfo_colptrs(n, nsuper, sstart, upd_ptr, upd_idx, Lp, Up);
    i64 *Lcur = (i64 *)malloc((size_t)(n > 0 ? n : 1) * sizeof(i64));
    i64 *Ucur = (i64 *)malloc((size_t)(n > 0 ? n : 1) * sizeof(i64));
    memcpy(Lcur, Lp, (size_t)n * sizeof(i64));
    memcpy(Ucur, Up, (size_t)n * sizeof(i64));
    int status = 0;

    for (i64 s = 0; s < nsuper; s++) {
        i64 c1 = sstart[s], c2 = sstart[s + 1] - 1, np = c2 - c1 + 1;
        const i64 *upd = upd_idx + upd_ptr[s];
        i64 nu = upd_ptr[s + 1] - upd_ptr[s], nf = np + nu, ld = nf;
        for (i64 a = 0; a < np; a++) loc[c1 + a] = a;
        for (i64 t = 0; t < nu; t++) loc[upd[t]] = np + t;
        double *F = Fbuf;
        memset(F, 0, (size_t)(nf * nf) * sizeof(double));

        for (i64 a = 0; a < np; a++) {
            i64 j = c1 + a;
            for (i64 p = Vp[j]; p < Vp[j + 1]; p++) {
                i64 i = Vi[p];
                if (i >= j && loc[i] >= 0) F[loc[i] + a * ld] += Vx[p];
            }
            for (i64 p = Tp[j]; p < Tp[j + 1]; p++) {
                i64 k = Ti[p];
                if (k > j && loc[k] >= 0) F[a + loc[k] * ld] += Tx[p];
            }
        }

        for (i64 ci = child_ptr[s + 1] - 1; ci >= child_ptr[s]; ci--) {
            i64 ck = child_idx[ci];
            const i64 *crows = upd_idx + upd_ptr[ck];
            i64 m = upd_ptr[ck + 1] - upd_ptr[ck];
            if (m == 0) continue;
            const double *C = cb + cboff[ck];
            for (i64 cs = 0; cs < m; cs++) {
                i64 dc = loc[crows[cs]];
                const double *col = C + cs * m;
                for (i64 r = 0; r < m; r++) F[loc[crows[r]] + dc * ld] += col[r];
            }
            cbtop = cboff[ck];
        }

        i64 info = 0;
        if (backend == 1 && p_getrf) {
            i64 npv = np, ldv = ld, inf = 0;
            p_getrf(&npv, &npv, F, &ldv, ipiv, &inf);
            for (i64 a = 0; a < np; a++) ipiv[a] -= 1;
            info = inf;
        } else {
            info = getf2_loops(np, F, ld, ipiv);
        }
        if (info != 0) {
            if (*info_col < 0) *info_col = c1 + info - 1;
            if (check) { status = 1; break; }
        }

        if (nu > 0) {
            double *R12 = F + np * ld;
            for (i64 k = 0; k < np; k++) {
                i64 pk = ipiv[k];
                if (pk != k)
                    for (i64 c = 0; c < nu; c++) {
                        double t = R12[k + c * ld]; R12[k + c * ld] = R12[pk + c * ld]; R12[pk + c * ld] = t;
                    }
            }
            double *L21 = F + np;
            double *F22 = F + np + np * ld;
            if (backend == 1 && p_trsm && p_gemm) {
                const double one = 1.0, mone = -1.0;
                i64 nuv = nu, npv = np, ldv = ld;
                p_trsm("R", "U", "N", "N", &nuv, &npv, &one, F, &ldv, L21, &ldv, 1, 1, 1, 1);
                p_trsm("L", "L", "N", "U", &npv, &nuv, &one, F, &ldv, R12, &ldv, 1, 1, 1, 1);
                p_gemm("N", "N", &nuv, &nuv, &npv, &mone, L21, &ldv, R12, &ldv, &one, F22, &ldv, 1, 1);
            } else {
                trsm_run_loops(nu, np, F, ld, L21, ld);
                trsm_llu_loops(np, nu, F, ld, R12, ld);
                gemm_loops(nu, nu, np, L21, ld, R12, ld, F22, ld);
            }
        }

        for (i64 a = 0; a < np; a++) orig[a] = c1 + a;
        for (i64 k = 0; k < np; k++) {
            i64 pk = ipiv[k];
            if (pk != k) { i64 t = orig[k]; orig[k] = orig[pk]; orig[pk] = t; }
        }
        for (i64 a = 0; a < np; a++) prow[c1 + a] = orig[a];

        for (i64 a = 0; a < np; a++) {
            i64 g = c1 + a, lp = Lcur[g];
            Li[lp] = orig[a]; Lx[lp] = 1.0; lp++;
            for (i64 b = a + 1; b < np; b++) { Li[lp] = orig[b]; Lx[lp] = F[b + a * ld]; lp++; }
            for (i64 t = 0; t < nu; t++) { Li[lp] = upd[t]; Lx[lp] = F[np + t + a * ld]; lp++; }
            Lcur[g] = lp;
            i64 up = Ucur[g];
            for (i64 b = 0; b <= a; b++) { Ui[up] = c1 + b; Ux[up] = F[b + a * ld]; up++; }
            Ucur[g] = up;
        }
        for (i64 t = 0; t < nu; t++) {
            i64 J = upd[t], up = Ucur[J];
            for (i64 a = 0; a < np; a++) { Ui[up] = c1 + a; Ux[up] = F[a + (np + t) * ld]; up++; }
            Ucur[J] = up;
        }

        if (nu > 0) {
            if (cbtop + nu * nu > cbcap) {
                i64 nc = cbtop + nu * nu;
                if (nc < 2 * cbcap) nc = 2 * cbcap;
                cb = (double *)realloc(cb, (size_t)nc * sizeof(double));
                cbcap = nc;
            }
            cboff[s] = cbtop;
            for (i64 c = 0; c < nu; c++)
                memcpy(cb + cbtop + c * nu, F + np + (np + c) * ld, (size_t)nu * sizeof(double));
            cbtop += nu * nu;
        }
        for (i64 a = 0; a < np; a++) loc[c1 + a] = -1;
        for (i64 t = 0; t < nu; t++) loc[upd[t]] = -1;
    }

    if (status == 0) {
        i64 *rowfac = (i64 *)malloc((size_t)(n > 0 ? n : 1) * sizeof(i64));
        for (i64 k = 0; k < n; k++) rowfac[prow[k]] = k;
        for (i64 p = 0; p < Lp[n]; p++) Li[p] = rowfac[Li[p]];
        free(rowfac);
    }
    free(Tp); free(Ti); free(Tx); free(loc); free(Fbuf); free(cb); free(cboff); free(ipiv);
    free(orig); free(Lcur); free(Ucur);
    return status;
}

/* ---------------- solves (src/solve.jl, src/interface.jl) ---------------- */

void mfo_lsolve(i64 n, const i64 *Lp, const i64 *Li, const double *Lx, double *y)
{
    for (i64 j = 0; j < n; j++) {
        double yj = y[j];
        for (i64 p = Lp[j] + 1; p < Lp[j + 1]; p++) y[Li[p]] -= Lx[p] * yj;
    }
}

void mfo_usolve(i64 n, const i64 *Up, const i64 *Ui, const double *Ux, double *y)
{
    for (i64 j = n - 1; j >= 0; j--) {
        i64 pd = Up[j + 1] - 1;
        double yj = y[j] / Ux[pd];
        y[j] = yj;
        for (i64 p = Up[j]; p < pd; p++) y[Ui[p]] -= Ux[p] * yj;
    }
}

/* x = Q U^-1 L^-1 P (Rs .* b); Rs == NULL means no scaling. work has n entries. */
void mfo_solve_factor(i64 n, const i64 *Lp, const i64 *Li, const double *Lx, const i64 *Up,
                      const i64 *Ui, const double *Ux, const i64 *p, const i64 *q,
                      const double *Rs, const double *b, double *x, double *work)
{
    for (i64 k = 0; k < n; k++) work[k] = Rs ? Rs[p[k]] * b[p[k]] : b[p[k]];
    mfo_lsolve(n, Lp, Li, Lx, work);
    mfo_usolve(n, Up, Ui, Ux, work);
    for (i64 k = 0; k < n; k++) x[q[k]] = work[k];
}

/* r = b - A x, A in CSC (column-oriented accumulation as SparseArrays' A*x). */
void mfo_residual(i64 n, const i64 *Ap, const i64 *Ai, const double *Ax, const double *x,
                  const double *b, double *r, double *work)
{
    for (i64 i = 0; i < n; i++) work[i] = 0.0;
    for (i64 j = 0; j < n; j++) {
        double xj = x[j];
        for (i64 p = Ap[j]; p < Ap[j + 1]; p++) work[Ai[p]] += Ax[p] * xj;
    }
    for (i64 i = 0; i < n; i++) r[i] = b[i] - work[i];
}
