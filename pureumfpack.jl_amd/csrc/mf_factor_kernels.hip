// mf_factor_kernels.hip -- numeric factorization kernels (assembly, extend-add, getrf,
// trsm, Schur gemm) for gfx950.  See mf_kernels.h for the front storage layout.
#include "mf_kernels.h"

namespace mf {

// ------------------------------------------------------------------------------------------
// Small fronts (nf <= small_nf): one workgroup per front, the whole front in LDS.
// Steps: zero the front, place the (Rs.*A)[qf,qf] entries (src/multifrontal.jl:216-231),
// extend-add the children's Schur blocks in reverse child order (:235-243, :45-57), then
// right-looking elimination over all nf rows with the pivot search confined to the np pivot
// rows (:247-252).  Eliminating the L21 rows and the R12 columns in the same sweep gives
// L21 = A21 U11^-1, U12 = L11^-1 P A12 and F22 - L21 U12 (:254-268).  The panel and U12 go
// to the factor store, F22 to this front's contribution block in the arena (ld nu).
// ------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_fused_small(FactorArgs a)
{
    extern __shared__ unsigned char smem_raw[];
    T *F = reinterpret_cast<T *>(smem_raw);
    const int mat = blockIdx.y;
    if (skip_matrix(a.only, mat)) return;
    const SnDev d = a.sn[a.list[blockIdx.x]];
    const int nf = d.nf, np = d.np, nu = d.nu;
    int *orig = reinterpret_cast<int *>(F + nf * nf);
    __shared__ int s_piv;
    const int tid = threadIdx.x, nt = blockDim.x;
    T *ws = reinterpret_cast<T *>(a.ws) + mat * a.ws_stride;
    T *store = reinterpret_cast<T *>(a.store) + mat * a.store_stride;
    const double *ax = a.ax + mat * a.ax_stride;
    const double *rs = a.rs ? a.rs + mat * a.rs_stride : nullptr;

    for (int i = tid; i < nf * nf; i += nt) F[i] = T(0);
    for (int i = tid; i < np; i += nt) orig[i] = d.c1 + i;
    __syncthreads();
    for (int e = tid; e < d.norig; e += nt) {
        int q = d.orig_off + e;
        F[a.orig_row[q] + a.orig_col[q] * nf] = scaled_value<T>(ax, rs, a.arow, a.orig_src[q]);
    }
    __syncthreads();
    for (int ci = 0; ci < d.nchild; ci++) {
        const SnDev c = a.sn[a.child_list[d.child_off + ci]];
        const int m = c.nu;
        const T *cb = ws + c.cboff;
        const int32_t *rm = a.rmap + c.upd_off;
        for (int idx = tid; idx < m * m; idx += nt) {
            int r = idx % m, s = idx / m;
            F[rm[r] + rm[s] * nf] += cb[r + (int64_t)s * c.cbld];
        }
        __syncthreads();
    }

    for (int k = 0; k < np; k++) {
        if (tid < 64) {
            T best = T(-1);
            int bi = 0x7fffffff;
            for (int r = k + tid; r < np; r += 64) {
                T v = fabs(F[r + k * nf]);
                if (v > best) { best = v; bi = r; }
            }
            wave_argmax(best, bi);
            if (tid == 0) s_piv = bi;
        }
        __syncthreads();
        const int p = s_piv;
        if (p != k) {
            for (int c = tid; c < nf; c += nt) {
                T t = F[k + c * nf];
                F[k + c * nf] = F[p + c * nf];
                F[p + c * nf] = t;
            }
            if (tid == 0) { int t = orig[k]; orig[k] = orig[p]; orig[p] = t; }
        }
        __syncthreads();
        const T piv = F[k + k * nf];
        if (piv == T(0) && tid == 0) atomicMin(a.info + mat, d.c1 + k);
        const T rcp = T(1) / piv;
        const bool small_piv = fabs(piv) < Tiny<T>::v();
        for (int r = k + 1 + tid; r < nf; r += nt) {
            T v = F[r + k * nf];
            if (r < np) {
                // LAPACK getf2: scale by the reciprocal unless the pivot is zero or tiny
                if (piv != T(0)) v = small_piv ? v / piv : v * rcp;
            } else {
                v = v * rcp;  // trsm R/U/N/N multiplies by 1/U(j,j)
            }
            F[r + k * nf] = v;
        }
        __syncthreads();
        const int w = nf - k - 1;
        for (int idx = tid; idx < w * w; idx += nt) {
            int r = k + 1 + idx % w, c = k + 1 + idx / w;
            F[r + c * nf] -= F[r + k * nf] * F[k + c * nf];
        }
        __syncthreads();
    }
    T *P = store + d.foff;             // [L11\U11; L21]: the first np columns, ld nf
    for (int i = tid; i < nf * np; i += nt) P[i] = F[i];
    T *U = store + d.u12off;           // U12, ld np
    for (int i = tid; i < np * nu; i += nt) U[i] = F[(i % np) + (np + i / np) * nf];
    T *CB = ws + d.cboff;              // F22, ld nu
    for (int i = tid; i < nu * nu; i += nt) CB[i] = F[(np + i % nu) + (np + i / nu) * nf];
    int32_t *prow = a.prow + mat * a.n_stride;
    for (int i = tid; i < np; i += nt) prow[d.c1 + i] = orig[i];
}

// ------------------------------------------------------------------------------------------
// Large fronts, step 1: assemble one front column per workgroup in LDS, then store it.
// The column receives its original entries, then each contributing child column in reverse
// child order (the extend-add of src/multifrontal.jl:45-57 split by parent column, so no two
// workgroups ever write the same element and the result does not depend on scheduling).
// ------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_asm_cols(FactorArgs a)
{
    extern __shared__ unsigned char smem_raw[];
    T *col = reinterpret_cast<T *>(smem_raw);
    const int mat = blockIdx.y;
    if (skip_matrix(a.only, mat)) return;
    const int2 w = reinterpret_cast<const int2 *>(a.list)[blockIdx.x];
    const SnDev d = a.sn[w.x];
    const int c = w.y, nf = d.nf;
    const int tid = threadIdx.x, nt = blockDim.x;
    T *ws = reinterpret_cast<T *>(a.ws) + mat * a.ws_stride;
    const double *ax = a.ax + mat * a.ax_stride;
    const double *rs = a.rs ? a.rs + mat * a.rs_stride : nullptr;

    for (int i = tid; i < nf; i += nt) col[i] = T(0);
    __syncthreads();
    const int fc = d.fcol + c;
    for (int q = a.orig_colptr[fc] + tid; q < a.orig_colptr[fc + 1]; q += nt)
        col[a.orig_row[q]] = scaled_value<T>(ax, rs, a.arow, a.orig_src[q]);
    __syncthreads();
    // child contributions in list order, four at a time: their records and first rows are
    // loaded together, then added one contribution per barrier (rows of two children may meet)
    const int q1 = a.ext_colptr[fc + 1];
    for (int q = a.ext_colptr[fc]; q < q1; q += 4) {
        ExtRec e[4];
        T v[4];
        int32_t ri[4];
#pragma unroll
        for (int j = 0; j < 4; j++) e[j] = q + j < q1 ? a.ext[q + j] : ExtRec{0, 0, 0};
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const bool on = tid < e[j].nu;
            v[j] = on ? ws[e[j].src + tid] : T(0);
            ri[j] = on ? a.rmap[e[j].roff + tid] : 0;
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (q + j < q1) {   // the same for every thread
                if (tid < e[j].nu) col[ri[j]] += v[j];
                // the remaining rows eight per thread at a time: all loads before the adds
                for (int r0 = tid + nt; r0 < e[j].nu; r0 += 8 * nt) {
                    T vv[8];
                    int32_t rr[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        const int r = r0 + u * nt;
                        const bool on = r < e[j].nu;
                        vv[u] = on ? ws[e[j].src + r] : T(0);
                        rr[u] = on ? a.rmap[e[j].roff + r] : -1;
                    }
#pragma unroll
                    for (int u = 0; u < 8; u++)
                        if (rr[u] >= 0) col[rr[u]] += vv[u];
                }
                __syncthreads();
            }
        }
    }
    T *dst = ws + d.woff + (int64_t)c * nf;
    for (int i = tid; i < nf; i += nt) dst[i] = col[i];
    if (c < d.np && tid == 0) a.prow[mat * a.n_stride + d.c1 + c] = d.c1 + c;
}

// ------------------------------------------------------------------------------------------
// Large fronts, step 2: factor the panel of pivot columns [k, k+kb) over the candidate rows
// [k, np) held in registers (RPT rows per thread), pivot by workgroup argmax (first maximum),
// then
//   * apply the kb row swaps to every column outside the panel (LAPACK dlaswp on A11 and the
//     R12 swaps of src/multifrontal.jl:256-263),
//   * U rows:  A[k:k+kb, c] <- L_kk^-1 A[k:k+kb, c] for c >= k+kb     (trsm L/L/N/U),
//   * L21 cols: A[r, k:k+kb] <- A[r, k:k+kb] U_kk^-1 for r >= np        (trsm R/U/N/N).
// The trailing update is a grouped gemm launched next.
// ------------------------------------------------------------------------------------------
template <typename T, int NT, int RPT>
__global__ __launch_bounds__(NT) void k_panel(FactorArgs a, int k)
{
    constexpr int KB = PANEL_KB;
    constexpr int NW = NT / 64;
    const int mat = blockIdx.y;
    if (skip_matrix(a.only, mat)) return;
    const SnDev d = a.sn[a.list[blockIdx.x]];
    const int np = d.np, nf = d.nf;
    const int kb = min(KB, np - k);
    const int m = np - k;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wid = tid >> 6;
    T *F = reinterpret_cast<T *>(a.ws) + mat * a.ws_stride + d.woff;

    __shared__ T s_redv[2][NW];
    __shared__ int s_redi[2][NW];
    __shared__ T s_prow[2][KB];
    __shared__ T s_jrow[2][KB];
    __shared__ int s_ipiv[KB];

    T x[RPT][KB];
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        const int i = tid + NT * r;
#pragma unroll
        for (int j = 0; j < KB; j++)
            x[r][j] = (i < m && j < kb) ? F[(k + i) + (int64_t)(k + j) * nf] : T(0);
    }

#pragma unroll
    for (int j = 0; j < KB; j++) {
        if (j >= kb) continue;  // kb is uniform over the workgroup
        const int par = j & 1;
        T best = T(-1);
        int bi = 0x7fffffff;
#pragma unroll
        for (int r = 0; r < RPT; r++) {
            const int i = tid + NT * r;
            if (i >= j && i < m) {
                T v = fabs(x[r][j]);
                if (v > best) { best = v; bi = i; }
            }
        }
        wave_argmax(best, bi);
        if (lane == 0) { s_redv[par][wid] = best; s_redi[par][wid] = bi; }
        __syncthreads();
        best = s_redv[par][0];
        bi = s_redi[par][0];
#pragma unroll
        for (int w2 = 1; w2 < NW; w2++) {
            T ob = s_redv[par][w2];
            int oi = s_redi[par][w2];
            if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
        }
        const int piv = bi;
#pragma unroll
        for (int r = 0; r < RPT; r++) {
            const int i = tid + NT * r;
            if (i == piv) {
#pragma unroll
                for (int jj = 0; jj < KB; jj++) s_prow[par][jj] = x[r][jj];
            }
            if (i == j) {
#pragma unroll
                for (int jj = 0; jj < KB; jj++) s_jrow[par][jj] = x[r][jj];
            }
        }
        __syncthreads();
        if (piv != j) {
#pragma unroll
            for (int r = 0; r < RPT; r++) {
                const int i = tid + NT * r;
                if (i == j) {
#pragma unroll
                    for (int jj = 0; jj < KB; jj++) x[r][jj] = s_prow[par][jj];
                }
                if (i == piv) {
#pragma unroll
                    for (int jj = 0; jj < KB; jj++) x[r][jj] = s_jrow[par][jj];
                }
            }
        }
        const T pv = s_prow[par][j];
        if (tid == 0) {
            s_ipiv[j] = piv;
            if (pv == T(0)) atomicMin(a.info + mat, d.c1 + k + j);
        }
        const T rcp = T(1) / pv;
        const bool small_piv = fabs(pv) < Tiny<T>::v();
#pragma unroll
        for (int r = 0; r < RPT; r++) {
            const int i = tid + NT * r;
            if (i > j && i < m) {
                T l = x[r][j];
                if (pv != T(0)) l = small_piv ? l / pv : l * rcp;
                x[r][j] = l;
#pragma unroll
                for (int jj = 0; jj < KB; jj++)
                    if (jj > j) x[r][jj] -= l * s_prow[par][jj];
            }
        }
    }

    // store the factored panel
#pragma unroll
    for (int r = 0; r < RPT; r++) {
        const int i = tid + NT * r;
        if (i < m) {
#pragma unroll
            for (int j = 0; j < KB; j++)
                if (j < kb) F[(k + i) + (int64_t)(k + j) * nf] = x[r][j];
        }
    }
    // The kb sequential swaps only move rows inside the set R = {0..kb-1} U {ipiv[j]} (panel
    // coordinates).  Compose them once: row dst[q] receives the old contents of row src[q].
    // k_panel_apply applies them to the other columns.
    if (tid == 0) {
        int dst[2 * KB], src[2 * KB], cont[2 * KB];
        int cnt = kb;
        for (int q = 0; q < kb; q++) dst[q] = q;
        for (int j = 0; j < kb; j++) {
            int p = s_ipiv[j], found = 0;
            for (int q = 0; q < cnt; q++) found |= (dst[q] == p);
            if (!found) dst[cnt++] = p;
        }
        for (int q = 0; q < cnt; q++) cont[q] = q;
        for (int j = 0; j < kb; j++) {
            int p = s_ipiv[j], qp = 0;
            for (int q = 0; q < cnt; q++)
                if (dst[q] == p) qp = q;
            int t = cont[j]; cont[j] = cont[qp]; cont[qp] = t;
        }
        for (int q = 0; q < cnt; q++) src[q] = dst[cont[q]];
        int32_t *pi = a.pinfo + (int64_t)mat * a.pinfo_stride + (int64_t)a.list[blockIdx.x] * PINFO_STRIDE;
        pi[0] = cnt;
        for (int q = 0; q < cnt; q++) { pi[1 + q] = dst[q]; pi[1 + 2 * KB + q] = src[q]; }
        int32_t *prow = a.prow + mat * a.n_stride + d.c1 + k;
        int orig[2 * KB];
        for (int q = 0; q < cnt; q++) orig[q] = prow[src[q]];
        for (int q = 0; q < cnt; q++) prow[dst[q]] = orig[q];
    }
}

// ------------------------------------------------------------------------------------------
// Large fronts, step 2 for pivot blocks taller than registers hold (np - k > 3072 rows): the
// same panel factorization with the panel kept in the front (global memory, L2-resident:
// kb <= 16 columns) instead of registers.  Thread t owns rows t, t + NT, ...; at step j the
// workgroup argmax (first maximum) picks the pivot row, the pivot row and row j are staged in
// LDS, and each thread rewrites only its own rows (the pivot row's owner writes old row j's
// values there, row j's owner the pivot row), so no two threads write one element.  Barriers
// order the steps; the global stores of one step are visible to the workgroup after them.
// ------------------------------------------------------------------------------------------
template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_panel_big(FactorArgs a, int k)
{
    constexpr int KB = PANEL_KB;
    constexpr int NW = NT / 64;
    const int mat = blockIdx.y;
    if (skip_matrix(a.only, mat)) return;
    const SnDev d = a.sn[a.list[blockIdx.x]];
    const int np = d.np, nf = d.nf;
    const int kb = min(KB, np - k);
    const int m = np - k;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wid = tid >> 6;
    T *F = reinterpret_cast<T *>(a.ws) + mat * a.ws_stride + d.woff + k + (int64_t)k * nf;   // panel origin

    __shared__ T s_redv[NW];
    __shared__ int s_redi[NW];
    __shared__ T s_prow[KB];
    __shared__ T s_jrow[KB];
    __shared__ int s_ipiv[KB];
    __shared__ int s_piv;

    for (int j = 0; j < kb; j++) {
        T best = T(-1);
        int bi = 0x7fffffff;
        for (int i = j + tid; i < m; i += NT) {
            const T v = fabs(F[i + (int64_t)j * nf]);
            if (v > best) { best = v; bi = i; }
        }
        wave_argmax(best, bi);
        if (lane == 0) { s_redv[wid] = best; s_redi[wid] = bi; }
        __syncthreads();
        if (tid == 0) {
            T b = s_redv[0];
            int p = s_redi[0];
            for (int w = 1; w < NW; w++) {
                const T ob = s_redv[w];
                const int oi = s_redi[w];
                if (ob > b || (ob == b && oi < p)) { b = ob; p = oi; }
            }
            s_piv = p;
        }
        __syncthreads();
        const int piv = s_piv;
        if (tid < kb) {
            s_prow[tid] = F[piv + (int64_t)tid * nf];
            s_jrow[tid] = F[j + (int64_t)tid * nf];
        }
        __syncthreads();
        const T pv = s_prow[j];
        if (tid == 0) {
            s_ipiv[j] = piv;
            if (pv == T(0)) atomicMin(a.info + mat, d.c1 + k + j);
        }
        const T rcp = T(1) / pv;
        const bool small_piv = fabs(pv) < Tiny<T>::v();
        for (int i = j + tid; i < m; i += NT) {
            if (i == j) {
                if (piv != j)
                    for (int jj = 0; jj < kb; jj++) F[i + (int64_t)jj * nf] = s_prow[jj];
                continue;
            }
            T l = i == piv ? s_jrow[j] : F[i + (int64_t)j * nf];
            if (pv != T(0)) l = small_piv ? l / pv : l * rcp;
            if (i == piv)
                for (int jj = 0; jj < j; jj++) F[i + (int64_t)jj * nf] = s_jrow[jj];
            F[i + (int64_t)j * nf] = l;
            for (int jj = j + 1; jj < kb; jj++) {
                const T v = i == piv ? s_jrow[jj] : F[i + (int64_t)jj * nf];
                F[i + (int64_t)jj * nf] = v - l * s_prow[jj];
            }
        }
        __syncthreads();
    }
    // compose the kb sequential swaps, as k_panel does
    if (tid == 0) {
        int dst[2 * KB], src[2 * KB], cont[2 * KB];
        int cnt = kb;
        for (int q = 0; q < kb; q++) dst[q] = q;
        for (int j = 0; j < kb; j++) {
            int p = s_ipiv[j], found = 0;
            for (int q = 0; q < cnt; q++) found |= (dst[q] == p);
            if (!found) dst[cnt++] = p;
        }
        for (int q = 0; q < cnt; q++) cont[q] = q;
        for (int j = 0; j < kb; j++) {
            int p = s_ipiv[j], qp = 0;
            for (int q = 0; q < cnt; q++)
                if (dst[q] == p) qp = q;
            int t = cont[j]; cont[j] = cont[qp]; cont[qp] = t;
        }
        for (int q = 0; q < cnt; q++) src[q] = dst[cont[q]];
        int32_t *pi = a.pinfo + (int64_t)mat * a.pinfo_stride + (int64_t)a.list[blockIdx.x] * PINFO_STRIDE;
        pi[0] = cnt;
        for (int q = 0; q < cnt; q++) { pi[1 + q] = dst[q]; pi[1 + 2 * KB + q] = src[q]; }
        int32_t *prow = a.prow + mat * a.n_stride + d.c1 + k;
        int orig[2 * KB];
        for (int q = 0; q < cnt; q++) orig[q] = prow[src[q]];
        for (int q = 0; q < cnt; q++) prow[dst[q]] = orig[q];
    }
}

// ------------------------------------------------------------------------------------------
// Large fronts, step 2b: after k_panel, for every front of the launch
//   * work item kind 0 (256 columns per workgroup, one column per thread): permute the rows
//     touched by the panel's swaps in every column outside the panel (LAPACK dlaswp on A11 and
//     the R12 swaps of src/multifrontal.jl:256-263), and for columns right of the panel solve
//     the U rows A[k:k+kb, c] <- L_kk^-1 A[k:k+kb, c]                (trsm L/L/N/U);
//   * work item kind 1 (256 rows per workgroup): L21 columns of the panel,
//     A[r, k:k+kb] <- A[r, k:k+kb] U_kk^-1 for r >= np               (trsm R/U/N/N).
// Items: int4 {front, kind, start, 0}.
// ------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_panel_apply(FactorArgs a, int k)
{
    constexpr int KB = PANEL_KB;
    __shared__ T s_blk[KB * KB];
    __shared__ int s_dst[2 * KB], s_src[2 * KB], s_cnt;
    const int mat = blockIdx.y;
    if (skip_matrix(a.only, mat)) return;
    const int4 w = reinterpret_cast<const int4 *>(a.list)[blockIdx.x];
    const SnDev d = a.sn[w.x];
    const int np = d.np, nf = d.nf;
    const int kb = min(KB, np - k);
    const int tid = threadIdx.x;
    T *F = reinterpret_cast<T *>(a.ws) + mat * a.ws_stride + d.woff;
    for (int e = tid; e < KB * KB; e += blockDim.x) {
        const int i = e % KB, j = e / KB;
        s_blk[e] = (i < kb && j < kb) ? F[(k + i) + (int64_t)(k + j) * nf] : T(0);
    }
    if (tid == 0) {
        const int32_t *pi = a.pinfo + (int64_t)mat * a.pinfo_stride + (int64_t)w.x * PINFO_STRIDE;
        const int cnt = pi[0];
        s_cnt = cnt;
        for (int q = 0; q < cnt; q++) { s_dst[q] = pi[1 + q]; s_src[q] = pi[1 + 2 * KB + q]; }
    }
    __syncthreads();
    if (w.y == 0) {
        const int cnt = s_cnt;
        int c = w.z + tid;
        if (c >= k) c += kb;  // skip the panel columns
        if (c >= nf) return;
        T *colp = F + (int64_t)c * nf + k;
        T v[2 * KB];
#pragma unroll
        for (int q = 0; q < 2 * KB; q++) v[q] = q < cnt ? colp[s_src[q]] : T(0);
        if (c >= k + kb) {
#pragma unroll
            for (int j = 0; j < KB; j++)
#pragma unroll
                for (int i = 0; i < KB; i++)
                    if (i > j) v[i] -= s_blk[i + j * KB] * v[j];
        }
#pragma unroll
        for (int q = 0; q < 2 * KB; q++)
            if (q < cnt) colp[s_dst[q]] = v[q];
    } else {
        const int r = np + w.z + tid;
        if (r >= nf) return;
        T y[KB];
#pragma unroll
        for (int j = 0; j < KB; j++) y[j] = j < kb ? F[r + (int64_t)(k + j) * nf] : T(0);
#pragma unroll
        for (int j = 0; j < KB; j++) {
            if (j < kb) {
                T v = y[j];
#pragma unroll
                for (int i = 0; i < KB; i++)
                    if (i < j) v -= y[i] * s_blk[i + j * KB];
                y[j] = v * (T(1) / s_blk[j + j * KB]);
            }
        }
#pragma unroll
        for (int j = 0; j < KB; j++)
            if (j < kb) F[r + (int64_t)(k + j) * nf] = y[j];
    }
}

// ------------------------------------------------------------------------------------------
// Large fronts, speculative path.  Partial pivoting restricted to the pivot block chooses, at
// step j, the first row of maximum magnitude among rows j..np-1.  If for every j the updated
// diagonal a_jj satisfies |a_ij| <= |a_jj| / tau for all candidate rows i > j (tau = 1: the
// diagonal IS that first maximum), the factorization without row interchanges is the
// pivoted one.  These kernels factor a 64-column block at a time without interchanges and
// check that condition on every multiplier; a violation sets a per-matrix flag and the
// factorization is redone on the pivoted schedule (k_panel / k_panel_apply).
//   k_spec_diag: one 256-thread workgroup per front factors the kb x kb diagonal block in
//                registers (four waves of 16 columns; one wave for special pivots).
//   k_spec_trsm: many workgroups; item int4 {front, kind, start, slot}:
//     kind 0: rows r in [start, start+SPEC_ROWS) of [k+kb, nf): L[r, k:k+kb] = A[r, k:k+kb] U_kk^-1,
//             checking the multiplier bound on the candidate rows r < np;
//     kind 1: columns c in [start, start+SPEC_ROWS) of [k+kb, nf): U[k:k+kb, c] = L_kk^-1 A[k:k+kb, c];
//     slot: the front's position in the diagonal launch (its reciprocal pivots and special-pivot
//     flag in a.spec).
// ------------------------------------------------------------------------------------------
// The kb steps of the diagonal block factorization by one wave (see k_spec_diag), with getf2's
// rules per step: a normal pivot scales by its reciprocal, a zero pivot leaves its column
// unscaled (and is reported through info), a tiny one divides.
template <typename T, int KB>
__device__ __forceinline__ void spec_diag_steps(T (&x)[KB], int lane, int kb, T itau, int &bad, int &special,
                                                T &myrcp, int32_t *info, int infoc)
{
#pragma unroll
    for (int jc = 0; jc < KB / 32; jc++)
#pragma unroll
    for (int jj = 0; jj < 32; jj++) {
        const int j = 32 * jc + jj;
        if (j >= kb) continue;
        const T piv = lane_value(x[j], j);
        const bool act = lane > j && lane < kb;
        T v = x[j];
        bad |= (int)(act && fabs(v) > fabs(piv) * itau);
        const T rp = T(1) / piv;
        if (fabs(piv) >= Tiny<T>::v()) {
            v = v * rp;
        } else {
            special = 1;
            if (piv == T(0)) {
                if (lane == 0) atomicMin(info, infoc + j);
            } else {
                v = v / piv;
            }
        }
        myrcp = lane == j ? rp : myrcp;
        const T l = act ? v : T(0);
        x[j] = act ? v : x[j];
#pragma unroll
        for (int c = 0; c < KB; c++)
            if (c > j) x[c] = fma(-l, lane_value(x[c], j), x[c]);
    }
}

// Multiplier slots of the four-wave factorization: one LDS word per step and row, EMPTY until
// the owner wave stores the multiplier.  EMPTY is a signaling NaN, which no arithmetic result
// is (operations quiet a signaling NaN operand), so a stored multiplier never looks empty.
template <typename T> struct SlotBits;
template <> struct SlotBits<float> {
    typedef uint32_t U;
    static constexpr U EMPTY = 0x7fa00000u;
    static __device__ __forceinline__ U to(float v) { return __float_as_uint(v); }
    static __device__ __forceinline__ float from(U u) { return __uint_as_float(u); }
};
template <> struct SlotBits<double> {
    typedef unsigned long long U;
    static constexpr U EMPTY = 0x7ff4000000000000ull;
    static __device__ __forceinline__ U to(double v) { return (U)__double_as_longlong(v); }
    static __device__ __forceinline__ double from(U u) { return __longlong_as_double((long long)u); }
};

// Four waves on a block with normal pivots: wave w holds columns 16w .. 16w+15 of every row
// (lane = row).  The owner wave of steps 16w .. 16w+15 takes each pivot from lane j, checks and
// scales its multipliers, stores them in the step's slots (KB x KB words in LDS, EMPTY at entry)
// and updates its own later columns; a wave to its right applies step j as soon as its slot
// words are filled (each lane polls its own word), so no wave waits at a barrier and the owner
// runs ahead.  A poll that would exceed its bound (it cannot: each wave waits only on waves to
// its left) reports a multiplier violation, which sends the matrix to the pivoted schedule.
// Every element receives the same updates (explicit fused multiply-adds) in the same order as
// in spec_diag_steps, so with normal pivots the results are bit-identical to the one-wave form.
template <typename T, int KB>
__device__ __forceinline__ void spec_diag_steps4(T (&x)[KB / 4], int lane, int w, T itau, int &bad, int &special,
                                                 T &myrcp, typename SlotBits<T>::U *slots)
{
    typedef typename SlotBits<T>::U U;
    typedef __attribute__((address_space(3))) U lds_u;
    constexpr int CW = KB / 4;
    lds_u *sl = (lds_u *)slots;
    const int wu = __builtin_amdgcn_readfirstlane(w);
    int stuck = 0;
    // the column groups of the waves to the left, in order (16 steps unrolled per group)
#pragma unroll 1
    for (int ow = 0; ow < wu; ow++) {
#pragma unroll
        for (int jl = 0; jl < CW; jl++) {
            const int j = ow * CW + jl;
            U b = __atomic_load_n(sl + j * KB + lane, __ATOMIC_RELAXED);
            for (int it = 0; b == SlotBits<T>::EMPTY && !stuck; it++) {
                if (it > (1 << 16)) stuck = 1;
                __builtin_amdgcn_s_sleep(1);
                b = __atomic_load_n(sl + j * KB + lane, __ATOMIC_RELAXED);
            }
            const T l = stuck ? T(0) : SlotBits<T>::from(b);
#pragma unroll
            for (int c = 0; c < CW; c++) x[c] = fma(-l, lane_value(x[c], j), x[c]);
        }
    }
    bad |= stuck;
    // this wave's own group
#pragma unroll
    for (int jl = 0; jl < CW; jl++) {
        const int j = wu * CW + jl;
        const T piv = lane_value(x[jl], j);
        const bool act = lane > j;
        T v = x[jl];
        bad |= (int)(act && fabs(v) > fabs(piv) * itau);
        const T rp = T(1) / piv;
        special |= (int)!(fabs(piv) >= Tiny<T>::v());
        v = v * rp;
        myrcp = lane == j ? rp : myrcp;
        const T l = act ? v : T(0);
        x[jl] = act ? v : x[jl];
        __atomic_store_n(sl + j * KB + lane, SlotBits<T>::to(l), __ATOMIC_RELAXED);
#pragma unroll
        for (int c = jl + 1; c < CW; c++) x[c] = fma(-l, lane_value(x[c], j), x[c]);
    }
}

template <typename T, int KB>
__device__ __forceinline__ void spec_diag_front(const FactorArgs &a, int k, float tau, int slot,
                                                typename SlotBits<T>::U *slots)
{
    // One block per 256-thread workgroup.  One wave: lane i owns row i of the block in
    // registers; step j reads the pivot row from lane j with v_readlane (no memory round
    // trip).  Control flow depends only on the pivot, which is uniform: rows a step must not
    // change (finished rows, the pivot row) get multiplier 0.  A block is first factored by
    // four waves assuming normal pivots (spec_diag_steps4); a zero or tiny pivot sends it
    // through the one-wave form with getf2's rules.  The factored block
    // goes back to the front, where the trsm reads it; this front's slot of a.spec takes the
    // reciprocal pivots and a flag for a zero or tiny pivot (which getf2 treats apart: no
    // scaling, or a division).
    const int mat = blockIdx.y;
    const SnDev d = a.sn[a.list[slot]];
    const int np = d.np, nf = d.nf;
    const int kb = min(KB, np - k);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    T *F = reinterpret_cast<T *>(a.ws) + mat * a.ws_stride + d.woff;
    T *S = reinterpret_cast<T *>(a.spec) + mat * a.spec_stride + (int64_t)slot * SPEC_SLOT;
    const T itau = T(1) / T(tau);
    {
        // A partial block (kb < KB) is padded with the identity: the padded rows get zero
        // multipliers, the padded columns of the real rows stay zero and the padded pivots stay 1,
        // so the real entries receive exactly the updates of the one-wave form, which skips the
        // steps j >= kb.
        constexpr int CW = KB / 4;
        T xs[CW];
#pragma unroll
        for (int c = 0; c < CW; c++) {
            const int col = CW * w + c;
            xs[c] = (lane < kb && col < kb) ? F[(k + lane) + (int64_t)(k + col) * nf] : T(lane == col ? 1 : 0);
        }
        for (int e = threadIdx.x; e < KB * KB; e += 256) slots[e] = SlotBits<T>::EMPTY;
        __syncthreads();
        int bad = 0, special = 0;
        T myrcp = T(1);
        spec_diag_steps4<T, KB>(xs, lane, w, itau, bad, special, myrcp, slots);
        if (!__syncthreads_or(special)) {
            if (bad) atomicOr(a.flag + mat, 1);
            if (lane < kb) {
#pragma unroll
                for (int c = 0; c < CW; c++)
                    if (CW * w + c < kb) F[(k + lane) + (int64_t)(k + CW * w + c) * nf] = xs[c];
            }
            if ((lane >> 4) == w) S[lane] = myrcp;   // lane j's pivot belongs to wave j / 16 (1 when padded)
            if (threadIdx.x == 0) S[KB] = T(0);
            return;
        }
        // a zero or tiny pivot: wave 0 redoes the block below (F still holds the input)
    }
    if (w != 0) return;
    T x[KB];
#pragma unroll
    for (int c = 0; c < KB; c++) x[c] = (lane < kb && c < kb) ? F[(k + lane) + (int64_t)(k + c) * nf] : T(0);
    int bad = 0, special = 0;
    T myrcp = T(1);
    spec_diag_steps<T, KB>(x, lane, kb, itau, bad, special, myrcp, a.info + mat, d.c1 + k);
    if (bad) atomicOr(a.flag + mat, 1);
    if (lane < kb) {
#pragma unroll
        for (int c = 0; c < KB; c++)
            if (c < kb) F[(k + lane) + (int64_t)(k + c) * nf] = x[c];
    }
    S[lane] = lane < kb ? myrcp : T(1);
    if (lane == 0) S[KB] = special ? T(1) : T(0);
}

template <typename T, int KB>
__global__ __launch_bounds__(256) void k_spec_diag(FactorArgs a, int k, float tau)
{
    __shared__ typename SlotBits<T>::U slots[KB * KB];
    if (skip_matrix(a.only, blockIdx.y)) return;
    spec_diag_front<T, KB>(a, k, tau, blockIdx.x, slots);
}


// front entries (idx, k + c) / (k + c, idx) of matrix mat: base pointers of a row / column
template <typename T>
__device__ __forceinline__ T *F_row_base(const FactorArgs &a, int mat, const SnDev &d, int idx, int k)
{
    return reinterpret_cast<T *>(a.ws) + mat * a.ws_stride + d.woff + idx + (int64_t)k * d.nf;
}
template <typename T>
__device__ __forceinline__ T *F_col_base(const FactorArgs &a, int mat, const SnDev &d, int idx, int k)
{
    return reinterpret_cast<T *>(a.ws) + mat * a.ws_stride + d.woff + (int64_t)idx * d.nf + k;
}

// value of lane o of each group of 4 lanes (a DPP quad_perm broadcast; o a literal)
template <int O>
__device__ __forceinline__ float quad_lane(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), O * 0x55, 0xf, 0xf, false));
}
template <int O>
__device__ __forceinline__ double quad_lane(double v)
{
    const unsigned long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)u, O * 0x55, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), O * 0x55, 0xf, 0xf, false);
    return __longlong_as_double(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
template <typename T>
__device__ __forceinline__ T quad_lane(T v, int o)
{
    switch (o) {   // o is a constant in the unrolled steps
    case 0: return quad_lane<0>(v);
    case 1: return quad_lane<1>(v);
    case 2: return quad_lane<2>(v);
    default: return quad_lane<3>(v);
    }
}

// LDS layout of the trsm's triangle: M[j][c] (strictly c > j, zero elsewhere) at row j, group
// c % 4, position c / 4; a group holds 16 values and a pad of one 16-byte chunk, so that a lane's
// four 16-byte reads of a row hit distinct banks for the four groups and stay aligned.
template <typename T, int KB> struct TrsmLds {
    static constexpr int GS = KB / 4 + 16 / (int)sizeof(T);   // group stride (elements)
    static constexpr int RS = 4 * GS;                         // row stride
    static constexpr int ELEMS = KB * RS + 3 * KB;            // M, its diagonal, reciprocals + flag
};

// Right-looking substitution on one row (kind 0) or column (kind 1) of the block, shared by a
// group of 4 lanes: lane g of the group holds the entries c = 4 t + g (t = 0..15) in x[t].  Step
// j: the owner lane of column j finishes x_j (kind 0: multiplier check and scaling), the group
// takes it by a DPP broadcast, and every lane updates its entries c > j with row j of M, read
// from LDS one step ahead.  Each x_c receives its updates in the order j = 0, 1, ..., as in the
// row-by-row dot-product form.  dg: the diagonal of U (kind 0), rc: the reciprocal pivots.
// Normal pivots only (spec_solve_group_special covers the others).
template <typename T, int KB, bool LROWS>
__device__ __forceinline__ int spec_solve_group(T (&x)[KB / 4], const T *M, const T *dg, const T *rc, int g,
                                                bool cand, T itau)
{
    constexpr int NT = KB / 4, RS = TrsmLds<T, KB>::RS, VW = 16 / (int)sizeof(T);
    typedef T vec_t __attribute__((ext_vector_type(VW)));
    const T *Mg = M + g * TrsmLds<T, KB>::GS;
    int bad = 0;
    T cur[NT], nxt[NT];
    T ucur = T(0), unxt = T(0), rcur = T(0), rnxt = T(0);
#pragma unroll
    for (int t = 0; t < NT; t += VW) {
        const vec_t q = *reinterpret_cast<const vec_t *>(Mg + t);
#pragma unroll
        for (int e = 0; e < VW; e++) cur[t + e] = q[e];
    }
    if (LROWS) {
        ucur = dg[0];
        rcur = rc[0];
    }
#pragma unroll
    for (int jc = 0; jc < KB / 32; jc++)
#pragma unroll
    for (int jj = 0; jj < 32; jj++) {
        const int j = 32 * jc + jj, owner = j & 3, tj = j >> 2;
        if (j + 1 < KB) {   // the next step's operands go out before this step's arithmetic
            // entries t >= (j + 1) / 4 only (a loaded register that is never read can be
            // reused while its load is in flight, which makes the next write wait for it)
            const int t0 = (j + 1) >> 2, t1 = (t0 + VW - 1) / VW * VW;
#pragma unroll
            for (int t = 0; t < NT; t++)
                if (t >= t0 && t < t1) nxt[t] = Mg[(j + 1) * RS + t];
#pragma unroll
            for (int t = 0; t < NT; t += VW)
                if (t >= t1) {
                    const vec_t q = *reinterpret_cast<const vec_t *>(Mg + (j + 1) * RS + t);
#pragma unroll
                    for (int e = 0; e < VW; e++) nxt[t + e] = q[e];
                }
            if (LROWS) {
                unxt = dg[j + 1];
                rnxt = rc[j + 1];
            }
        }
        T v = x[tj];
        if (LROWS) {
            const T u = ucur;
            bad |= (int)(g == owner && cand && fabs(v) > fabs(u) * itau);
            v = v * rcur;   // reciprocal scaling (getf2, trsm R/U/N/N)
            if (g == owner) x[tj] = v;
        }
        v = quad_lane(v, owner);   // from the group's owner lane
        // entries c = 4 t + g > j: every lane for t > tj, lanes g > owner for t = tj (a select)
#pragma unroll
        for (int t = 0; t < NT; t++) {
            if (t > tj) {
                x[t] -= v * cur[t];
            } else if (t == tj) {
                const T y = x[t] - v * cur[t];
                x[t] = g > owner ? y : x[t];
            }
        }
#pragma unroll
        for (int t = 0; t < NT; t++) cur[t] = nxt[t];
        ucur = unxt;
        rcur = rnxt;
        // keep the schedule one step deep: loads hoisted further ahead only add registers
        __builtin_amdgcn_sched_barrier(0);
    }
    return bad;
}

// Kind 0 when some pivot of the block is zero or below the smallest normal number, where getf2
// leaves the column unscaled or divides instead of scaling by the reciprocal: the same steps as
// spec_solve_group, its operands read at their use (a compact form for a rare case).
template <typename T, int KB>
__device__ __forceinline__ int spec_solve_group_special(T (&x)[KB / 4], const T *M, const T *dg, const T *rc,
                                                        int g, bool cand, T itau)
{
    constexpr int RS = TrsmLds<T, KB>::RS;
    const T *Mg = M + g * TrsmLds<T, KB>::GS;
    int bad = 0;
#pragma unroll
    for (int jc = 0; jc < KB / 32; jc++)
#pragma unroll
    for (int jj = 0; jj < 32; jj++) {
        const int j = 32 * jc + jj, owner = j & 3, tj = j >> 2;
        T v = x[tj];
        const T u = dg[j];
        bad |= (int)(g == owner && cand && fabs(v) > fabs(u) * itau);
        if (u == T(0)) v = cand ? v : v * rc[j];
        else if (fabs(u) < Tiny<T>::v()) v = cand ? v / u : v * rc[j];
        else v = v * rc[j];
        if (g == owner) x[tj] = v;
        v = quad_lane(v, owner);
#pragma unroll
        for (int t = 0; t < KB / 4; t++)
            if (4 * t + 3 > j) {
                if (4 * t + g > j) x[t] -= v * Mg[j * RS + t];
            }
    }
    return bad;
}

template <typename T, int KB>
__global__ __launch_bounds__(4 * SPEC_ROWS) void k_spec_trsm(FactorArgs a, int k, float tau,
                                                             const T *__restrict__ spec)
{
    // SPEC_ROWS rows (kind 0) or columns (kind 1) per workgroup, 4 lanes each (see
    // spec_solve_group), M[j][c] = U(j, c) for kind 0 and L(c, j) for kind 1 (c > j, zero
    // elsewhere) staged in LDS (TrsmLds layout) from the factored diagonal block in the front,
    // read down its columns; the diagonal of U, the reciprocal pivots and the special-pivot
    // flag (from this front's slot of a.spec) follow it.
    constexpr int GS = TrsmLds<T, KB>::GS, RS = TrsmLds<T, KB>::RS;
    __shared__ __align__(16) T sm[TrsmLds<T, KB>::ELEMS];
    T *const dg = sm + KB * RS;
    T *const rc = dg + KB;
    const int mat = blockIdx.y;
    if (skip_matrix(a.only, mat)) return;
    const int4 w = reinterpret_cast<const int4 *>(a.list)[blockIdx.x];
    const SnDev d = a.sn[w.x];
    const int np = d.np, nf = d.nf;
    const int kb = min(KB, np - k);
    const bool lrows = w.y == 0;
    const int g = threadIdx.x & 3;   // lane of the group
    const int idx = k + kb + w.z + (threadIdx.x >> 2);        // row (kind 0) or column (kind 1) in the front
    const bool on = idx < nf;
    const bool cand = lrows && idx < np;
    const T itau = T(1) / T(tau);
    // kind 0 reads row idx across the block columns (stride nf); kind 1 reads column idx
    // down the block rows (contiguous).  These loads do not depend on the staged block: they
    // go out first.
    T *base = lrows ? F_row_base<T>(a, mat, d, idx, k) : F_col_base<T>(a, mat, d, idx, k);
    const int64_t stride = lrows ? nf : 1;
    T x[KB / 4];
#pragma unroll
    for (int t = 0; t < KB / 4; t++) {
        const int c = 4 * t + g;
        x[t] = (on && c < kb) ? base[c * stride] : T(0);
    }
    {
        const T *S = spec + mat * a.spec_stride + (int64_t)w.w * SPEC_SLOT;
        const T *Fd = reinterpret_cast<const T *>(a.ws) + mat * a.ws_stride + d.woff + k + (int64_t)k * nf;
        constexpr int PT = KB * KB / (4 * SPEC_ROWS);
        T v[PT];
#pragma unroll
        for (int h = 0; h < PT; h++) {
            const int e = threadIdx.x + h * 4 * SPEC_ROWS, i = e % KB, jj = e / KB;   // block entry (i, jj)
            v[h] = (i < kb && jj < kb) ? Fd[i + (int64_t)jj * nf] : T(0);
        }
        if (lrows) {
#pragma unroll
            for (int h = 0; h < PT; h++) {   // M[i][jj] = U(i, jj)
                const int e = threadIdx.x + h * 4 * SPEC_ROWS, i = e % KB, jj = e / KB;
                sm[i * RS + (jj & 3) * GS + (jj >> 2)] = jj > i ? v[h] : T(0);
                if (jj == i) dg[i] = v[h];
            }
        } else {
#pragma unroll
            for (int h = 0; h < PT; h++) {   // M[jj][i] = L(i, jj)
                const int e = threadIdx.x + h * 4 * SPEC_ROWS, i = e % KB, jj = e / KB;
                sm[jj * RS + (i & 3) * GS + (i >> 2)] = i > jj ? v[h] : T(0);
            }
        }
        for (int e = threadIdx.x; e < 2 * KB; e += 4 * SPEC_ROWS) rc[e] = S[e];
    }
    __syncthreads();
    int bad = 0;
    if (!lrows) {
        spec_solve_group<T, KB, false>(x, sm, dg, rc, g, false, itau);
    } else {
        if (rc[KB] == T(0))
            bad = spec_solve_group<T, KB, true>(x, sm, dg, rc, g, cand, itau);
        else
            bad = spec_solve_group_special<T, KB>(x, sm, dg, rc, g, cand, itau);
    }
    if (bad) atomicOr(a.flag + mat, 1);
    if (on) {
#pragma unroll
        for (int t = 0; t < KB / 4; t++) {
            const int c = 4 * t + g;
            if (c < kb) base[c * stride] = x[t];
        }
    }
}

// ------------------------------------------------------------------------------------------
// Inverses of the pivot-block factors of large fronts (np > INV_NP), used by the solves to turn
// the triangular sweeps into matrix-vector products.  Leaves: the 64x64 diagonal blocks,
// inverted by substitution on the identity, one column per thread; the off-diagonal blocks
// come from the recursive rule inv([A 0; M B]) = [inv(A) 0; -inv(B) M inv(A) inv(B)] as two
// batched gemms per tree level (and the transposed rule for U).  Items int4 {front, b0}.
// ------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(128) void k_inv_leaf(FactorArgs a)
{
    constexpr int B = INV_LEAF, LD = INV_LEAF + 1;
    __shared__ T sA[B * LD];
    __shared__ T sX[B * LD];
    const int mat = blockIdx.y;
    if (skip_matrix(a.only, mat)) return;
    const int4 w = reinterpret_cast<const int4 *>(a.list)[blockIdx.x];
    const SnDev d = a.sn[w.x];
    const int np = d.np, nf = d.nf, b0 = w.y;
    const int bs = min(B, np - b0);
    T *pool = reinterpret_cast<T *>(a.store) + mat * a.store_stride;
    const T *F = pool + d.foff;   // the stored panel (ld nf)
    const int tid = threadIdx.x;
    const bool upper = tid >= 64;          // threads 0..63: inv(L) columns, 64..127: inv(U) columns
    const int c = tid & 63;
    T *X = pool + d.ioff + (upper ? (int64_t)np * np : 0);
    for (int e = tid; e < B * B; e += blockDim.x) {
        const int i = e % B, j = e / B;
        sA[i + j * LD] = (i < bs && j < bs) ? F[(b0 + i) + (int64_t)(b0 + j) * nf] : T(0);
    }
    __syncthreads();
    // L part: unit lower (strictly-lower entries of sA); U part: upper incl. diagonal.
    if (!upper) {
        for (int i = 0; i < B; i++) sX[i + c * LD] = T(0);
        if (c < bs) {
            sX[c + c * LD] = T(1);
            for (int i = c + 1; i < bs; i++) {
                T acc = T(0);
                for (int q = c; q < i; q++) acc -= sA[i + q * LD] * sX[q + c * LD];
                sX[i + c * LD] = acc;
            }
        }
    }
    __syncthreads();
    if (!upper) {
        for (int e = c; e < bs * bs; e += 64) {
            const int i = e % bs, j = e / bs;
            X[(b0 + i) + (int64_t)(b0 + j) * np] = sX[i + j * LD];
        }
    }
    __syncthreads();
    if (upper) {
        for (int i = 0; i < B; i++) sX[i + c * LD] = T(0);
        if (c < bs) {
            sX[c + c * LD] = T(1) / sA[c + c * LD];
            for (int i = c - 1; i >= 0; i--) {
                T acc = T(0);
                for (int q = i + 1; q <= c; q++) acc -= sA[i + q * LD] * sX[q + c * LD];
                sX[i + c * LD] = acc / sA[i + i * LD];
            }
        }
    }
    __syncthreads();
    if (upper) {
        for (int e = c; e < bs * bs; e += 64) {
            const int i = e % bs, j = e / bs;
            X[(b0 + i) + (int64_t)(b0 + j) * np] = sX[i + j * LD];
        }
    }
}

// ------------------------------------------------------------------------------------------
// Grouped gemm C -= A B on MFMA (v_mfma_*_16x16x4): 64x64 tile per 256-thread workgroup,
// four waves in a 2x2 arrangement, each owning a 32x32 quadrant (2x2 MFMA blocks).  A and B
// tiles are staged through LDS k-major so the MFMA operand reads are conflict free.
// ------------------------------------------------------------------------------------------
template <typename T> struct Mfma;
// Output register q of lane l holds C[row(l, q)][l % 16].  The FP64 instruction interleaves the
// four lane groups row by row (row = 4q + l/16); the FP32 one gives each lane group four
// consecutive rows (row = 4(l/16) + q).
template <> struct Mfma<double> {
    typedef v4d acc_t;
    static __device__ __forceinline__ acc_t mma(double x, double y, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int q) { return 4 * q + (lane >> 4); }
};
template <> struct Mfma<float> {
    typedef v4f acc_t;
    static __device__ __forceinline__ acc_t mma(float x, float y, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(x, y, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int q) { return 4 * (lane >> 4) + q; }
};

// USE (0 panel update / inversion, 1 Schur complement, 2 Schur complement of a front with
// nf >= 1024) only names the instantiation, so that profilers report them apart; the code is
// the same.
//
// 4 waves in a 2 x 2 arrangement, each owning a (TM/2) x (TN/2) block of C as FM x FN MFMA
// 16 x 16 accumulators.  Slices of BK columns of A and rows of B are double-buffered in LDS:
// the global loads of slice s+1 are in flight (registers) while slice s is multiplied, then
// stored to the other buffer, one barrier per slice.  The MFMA computes B^T A^T, so each lane
// ends up with C elements of consecutive rows: the read-modify-write of C is coalesced.
// Accumulator block i of a wave covers the rows wm + FM s + i (s = lane % 16, the MFMA's
// row slot), interleaved rather than 16 apart, so a lane's FM operands of one k step are
// adjacent in LDS and arrive in one wide read (likewise FN for B).  Every element of C is
// accumulated over k in the same order (k steps of 4, ascending) for every tile shape, so the
// choice of tile never changes a result bit.
template <typename T> struct GemmCfg {
    static constexpr int BK = 16;                          // k per LDS slice
    static constexpr int PAD = sizeof(T) == 8 ? 2 : 4;    // row padding: keeps 16-byte rows, staggers banks
    static constexpr int V = 16 / sizeof(T);              // elements per 16-byte global load
};
// LDS elements a gemm tile needs (two buffers of the A and B slices)
template <typename T, int TM, int TN>
constexpr int gemm_lds_elems() { return 2 * GemmCfg<T>::BK * ((TM + GemmCfg<T>::PAD) + (TN + GemmCfg<T>::PAD)); }

// FM consecutive elements of LDS into registers (the compiler merges the loads into wide reads)
template <typename T, int F>
__device__ __forceinline__ void lds_run(const T *p, T (&v)[F])
{
#pragma unroll
    for (int i = 0; i < F; i++) v[i] = p[i];
}

// Global loads are 16 bytes wide (V elements): A slices as V consecutive rows of a column, B
// slices as V consecutive k of a column.  Fronts have any leading dimension, so these loads are
// only element-aligned (the GPU's unaligned access mode serves them); a slice that reaches past
// the problem's edge is loaded element by element with bounds checks instead.
template <typename T, int TM, int TN, int PF = 1>
__device__ __forceinline__ void gemm_tile(const GemmProb *__restrict__ probs, const GemmTile tl, void *pool_v,
                                          int64_t pool_stride, T *lds)
{
    typedef typename Mfma<T>::acc_t acc_t;
    constexpr int BK = GemmCfg<T>::BK, PAD = GemmCfg<T>::PAD, V = GemmCfg<T>::V;
    typedef T vec_t __attribute__((ext_vector_type(V), aligned(sizeof(T))));
    constexpr int LA = TM + PAD, LB = TN + PAD;
    constexpr int WM = TM / 2, WN = TN / 2, FM = WM / 16, FN = WN / 16;
    constexpr int AV = TM * BK / V / 256, BV = TN * BK / V / 256;
    static_assert(AV >= 1 && BV >= 1 && AV * V * 256 == TM * BK && BV * V * 256 == TN * BK, "tile and slice shapes");
    T *As = lds;                      // [2][BK][LA]
    T *Bs = lds + 2 * BK * LA;        // [2][BK][LB]
    const GemmProb pr = probs[tl.prob];
    T *pool = reinterpret_cast<T *>(pool_v) + blockIdx.y * pool_stride;
    const T *A = pool + pr.aoff;
    const T *B = pool + pr.boff;
    T *C = pool + pr.coff;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = (wid & 1) * WM, wn = (wid >> 1) * WN;
    const int m0 = tl.tm, n0 = tl.tn;
    const int slot = lane & 15, kl = lane >> 4;
    acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++) acc[i][j] = acc_t{0, 0, 0, 0};
    vec_t ra[PF][AV], rb[PF][BV];
    const bool mfull = m0 + TM <= pr.m, nfull = n0 + TN <= pr.n;
    auto gload = [&](int k0, vec_t (&ra)[AV], vec_t (&rb)[BV]) {
        const bool kfull = k0 + BK <= pr.k;
#pragma unroll
        for (int i = 0; i < AV; i++) {
            const int e = tid + 256 * i, rv = e % (TM / V), kk = e / (TM / V);
            const int gr = m0 + V * rv, gk = k0 + kk;
            const T *p = A + gr + (int64_t)gk * pr.lda;
            if (mfull && kfull) {
                ra[i] = *reinterpret_cast<const vec_t *>(p);
            } else {
#pragma unroll
                for (int q = 0; q < V; q++) ra[i][q] = (gr + q < pr.m && gk < pr.k) ? p[q] : T(0);
            }
        }
#pragma unroll
        for (int i = 0; i < BV; i++) {
            const int e = tid + 256 * i, kv = e % (BK / V), cc = e / (BK / V);
            const int gc = n0 + cc, gk = k0 + V * kv;
            const T *p = B + gk + (int64_t)gc * pr.ldb;
            if (nfull && kfull) {
                rb[i] = *reinterpret_cast<const vec_t *>(p);
            } else {
#pragma unroll
                for (int q = 0; q < V; q++) rb[i][q] = (gc < pr.n && gk + q < pr.k) ? p[q] : T(0);
            }
        }
    };
    auto sstore = [&](int buf, const vec_t (&ra)[AV], const vec_t (&rb)[BV]) {
#pragma unroll
        for (int i = 0; i < AV; i++) {
            const int e = tid + 256 * i, rv = e % (TM / V), kk = e / (TM / V);
            *reinterpret_cast<vec_t *>(As + (buf * BK + kk) * LA + V * rv) = ra[i];
        }
#pragma unroll
        for (int i = 0; i < BV; i++) {
            const int e = tid + 256 * i, kv = e % (BK / V), cc = e / (BK / V);
#pragma unroll
            for (int q = 0; q < V; q++) Bs[(buf * BK + V * kv + q) * LB + cc] = rb[i][q];
        }
    };
    // k range that can hold nonzero products for this tile (triangular operands); the slices
    // skipped contribute exact zeros, so the sums are unchanged
    int kbeg = 0, kend = pr.k;
    if (pr.flags & GEMM_A_LOWER) kend = min(kend, m0 + TM);
    if (pr.flags & GEMM_A_UPPER) kbeg = max(kbeg, m0);
    if (pr.flags & GEMM_B_LOWER) kbeg = max(kbeg, n0);
    if (pr.flags & GEMM_B_UPPER) kend = min(kend, n0 + TN);
    kbeg = kbeg / BK * BK;
    const int nk = kend > kbeg ? (kend - kbeg + BK - 1) / BK : 0;
    // C of a small tile is read before the k loop (its loads overlap the slices), not after it
    constexpr bool CPRE = FM * FN <= 4;
    const bool add = pr.flags & GEMM_ADD, set = pr.flags & GEMM_SET;
    const int gr0 = m0 + wm + FM * slot;
    T cpre[CPRE ? FN : 1][4][CPRE ? FM : 1];
    if (CPRE && !set) {
#pragma unroll
        for (int j = 0; j < FN; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int gc = n0 + wn + FN * Mfma<T>::row(lane, q) + j;
                const T *cp = C + (int64_t)gc * pr.ldc;
#pragma unroll
                for (int i = 0; i < FM; i++) cpre[j][q][i] = (gc < pr.n && gr0 + i < pr.m) ? cp[gr0 + i] : T(0);
            }
    }
    // PF slices are in flight in registers: slice s + PF is requested while slice s is multiplied,
    // so each load has PF slice times to arrive before it is stored to LDS
#pragma unroll
    for (int u = 0; u < PF; u++)
        if (u < nk) gload(kbeg + u * BK, ra[u], rb[u]);
    if (nk > 0) sstore(0, ra[0], rb[0]);
    __syncthreads();
    for (int s0 = 0; s0 < nk; s0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; u++) {
            const int s = s0 + u;
            if (s < nk) {
                const int cur = s & 1;
                // register set u held slice s, already in LDS: reuse it for slice s + PF
                if (s + PF < nk) gload(kbeg + (s + PF) * BK, ra[u], rb[u]);
                const T *Ac = As + (cur * BK + kl) * LA + wm + FM * slot;
                const T *Bc = Bs + (cur * BK + kl) * LB + wn + FN * slot;
#pragma unroll
                for (int kq = 0; kq < BK; kq += 4) {
                    T av[FM], bv[FN];
                    lds_run<T, FM>(Ac + kq * LA, av);
                    lds_run<T, FN>(Bc + kq * LB, bv);
#pragma unroll
                    for (int i = 0; i < FM; i++)
#pragma unroll
                        for (int j = 0; j < FN; j++) acc[i][j] = Mfma<T>::mma(bv[j], av[i], acc[i][j]);
                }
                if (s + 1 < nk) sstore(cur ^ 1, ra[(u + 1) % PF], rb[(u + 1) % PF]);
                __syncthreads();
            }
        }
    }
    // acc[i][j][q] = C[wm + FM (lane % 16) + i][wn + FN row(lane, q) + j]
#pragma unroll
    for (int j = 0; j < FN; j++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int gc = n0 + wn + FN * Mfma<T>::row(lane, q) + j;
            if (gc >= pr.n) continue;
            T *cp = C + (int64_t)gc * pr.ldc;
#pragma unroll
            for (int i = 0; i < FM; i++) {
                const int gr = gr0 + i;
                if (gr < pr.m) {
                    const T v = add ? acc[i][j][q] : -acc[i][j][q];
                    const T c0 = CPRE ? cpre[CPRE ? j : 0][q][CPRE ? i : 0] : cp[gr];
                    cp[gr] = set ? v : c0 + v;
                }
            }
        }
}

// slices in flight per tile shape: the 64 x 64 tile runs few MFMAs per slice, so its loads get
// more than one slice time to arrive (measured on the Schur launches of 3D Poisson 48^3: three
// slices in FP32, 2.5% less time than one; two in FP64, whose slices hold twice the registers,
// 4% less)
#ifndef MF_GEMM_PF
#define MF_GEMM_PF 3
#endif
#ifndef MF_GEMM_PF64
#define MF_GEMM_PF64 2
#endif
template <typename T, int TM, int TN> struct GemmPf {
    static constexpr int value = TM * TN <= 64 * 64 ? (sizeof(T) == 4 ? MF_GEMM_PF : MF_GEMM_PF64) : 1;
};

template <typename T, int TM, int TN, int USE>
__global__ __launch_bounds__(256) void k_gemm(const GemmProb *__restrict__ probs,
                                              const GemmTile *__restrict__ tiles, void *pool_v,
                                              int64_t pool_stride, const int32_t *__restrict__ only)
{
    __shared__ __align__(16) T lds[gemm_lds_elems<T, TM, TN>()];
    if (skip_matrix(only, blockIdx.y)) return;
    gemm_tile<T, TM, TN, GemmPf<T, TM, TN>::value>(probs, tiles[blockIdx.x], pool_v, pool_stride, lds);
}

// ------------------------------------------------------------------------------------------
// Schur complements in BF16 (precision MF_BF16, fronts with nf >= 1024): C -= A B with A, B
// and C in FP32 fronts.
//   k_pack_bf16: one workgroup per operand tile of PK_ROWS rows of A (or columns of B) by
//     PK_K k values: the tile is staged in LDS, its largest magnitude gives a power-of-two
//     exponent e (max |v| < 2^e), and v * 2^-e is rounded to BF16 (nearest even) into a
//     contiguous tile laid out [row][k] (k fastest), the order the MFMA operands are read in.
//   k_gemm_bf16: PK_ROWS x PK_ROWS of C per 256-thread workgroup (2 x 2 waves of 64 x 64, MFMA
//     16x16x32 BF16 with FP32 accumulation), one packed k-slice of A and B per step, double-
//     buffered in LDS with 16-byte loads and stores; the slice's FP32 partial product is added
//     to the FP32 result times 2^(ea + eb), which is exact.
// ------------------------------------------------------------------------------------------
typedef short bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16v8_t __attribute__((ext_vector_type(8)));

__device__ __forceinline__ unsigned short to_bf16(float x)
{
    unsigned u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even (finite inputs)
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// The exponent e of mx (mx in [2^(e-1), 2^e)), 0 for mx == 0.  e is kept >= -100 so that
// the scale 2^-e stays finite: fill entries of strongly convective operators decay
// exponentially and reach FP32 subnormals, where 2^-e would overflow (and 0 * inf = NaN).  A
// tile below 2^-101 then keeps its tiny magnitude, and 2^(ea + eb) flushes its product.
__device__ __forceinline__ int pow2_exp(float mx)
{
    if (!(mx > 0.0f) || !isfinite(mx)) return 0;
    int e;
    (void)frexpf(mx, &e);
    return max(e, -100);
}

// items int4 {problem, operand (0 A, 1 B), row tile, k tile}
__global__ __launch_bounds__(256) void k_pack_bf16(FactorArgs a, const GemmProb *__restrict__ probs)
{
    constexpr int LD = PK_K + 1;
    __shared__ float t[PK_ROWS * LD];
    __shared__ float smax[4];
    const int mat = blockIdx.y;
    if (skip_matrix(a.only, mat)) return;
    const int4 w = reinterpret_cast<const int4 *>(a.list)[blockIdx.x];
    const GemmProb pr = probs[w.x];
    const bool isA = w.y == 0;
    const float *ws = reinterpret_cast<const float *>(a.ws) + mat * a.ws_stride;
    const int tid = threadIdx.x, lane = tid & 63;
    const int r0 = w.z * PK_ROWS, k0 = w.w * PK_K;
    float mx = 0.0f;
    if (isA) {   // A(m x k), column-major: threads walk rows (coalesced)
        const float *A = ws + pr.aoff;
        const int r = tid % PK_ROWS;
#pragma unroll 4
        for (int kk = tid / PK_ROWS; kk < PK_K; kk += 256 / PK_ROWS) {
            const int gr = r0 + r, gk = k0 + kk;
            const float v = (gr < pr.m && gk < pr.k) ? A[gr + (int64_t)gk * pr.lda] : 0.0f;
            t[r * LD + kk] = v;
            mx = fmaxf(mx, fabsf(v));
        }
    } else {     // B(k x n), column-major: threads walk k (coalesced); tile row = column of B
        const float *B = ws + pr.boff;
        const int kk = tid % PK_K;
#pragma unroll 4
        for (int c = tid / PK_K; c < PK_ROWS; c += 256 / PK_K) {
            const int gc = r0 + c, gk = k0 + kk;
            const float v = (gc < pr.n && gk < pr.k) ? B[gk + (int64_t)gc * pr.ldb] : 0.0f;
            t[c * LD + kk] = v;
            mx = fmaxf(mx, fabsf(v));
        }
    }
    mx = wave_max(mx);
    if (lane == 0) smax[tid >> 6] = mx;
    __syncthreads();
    const int e = pow2_exp(fmaxf(fmaxf(smax[0], smax[1]), fmaxf(smax[2], smax[3])));
    const float sc = ldexpf(1.0f, -e);
    const int64_t tile = (isA ? pr.apk : pr.bpk) + (int64_t)w.z * ((pr.k + PK_K - 1) / PK_K) + w.w;
    bf16x8_t *dst = reinterpret_cast<bf16x8_t *>(reinterpret_cast<unsigned short *>(a.pack) +
                                                 (mat * a.pack_tiles + tile) * (int64_t)PK_TILE);
#pragma unroll
    for (int h = 0; h < PK_TILE / 8 / 256; h++) {
        const int ch = tid + 256 * h, row = ch / (PK_K / 8), kc = (ch % (PK_K / 8)) * 8;
        bf16x8_t v;
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = (short)to_bf16(t[row * LD + kc + q] * sc);
        dst[ch] = v;
    }
    if (tid == 0) a.pexp[mat * a.pack_tiles + tile] = e;
}

__global__ __launch_bounds__(256) void k_gemm_bf16(const GemmProb *__restrict__ probs,
                                                   const GemmTile *__restrict__ tiles, void *pool_v,
                                                   int64_t pool_stride, const int32_t *__restrict__ only,
                                                   const unsigned short *__restrict__ pack,
                                                   const int32_t *__restrict__ pexp, int64_t pack_tiles)
{
    constexpr int TM = PK_ROWS, BK = PK_K, WM = TM / 2, FM = WM / 16;   // 128, 64, 64, 4
    constexpr int LDK = BK + 8;                    // bf16 per LDS row (16-byte aligned, staggered banks)
    constexpr int CH = TM * BK / 8 / 256;          // 16-byte chunks per thread per operand tile
    __shared__ __align__(16) unsigned short As[2][TM * LDK];
    __shared__ __align__(16) unsigned short Bs[2][TM * LDK];
    typedef float acc4 __attribute__((ext_vector_type(4)));
    if (skip_matrix(only, blockIdx.y)) return;
    const GemmTile tl = tiles[blockIdx.x];
    const GemmProb pr = probs[tl.prob];
    const int mat = blockIdx.y;
    float *C = reinterpret_cast<float *>(pool_v) + mat * pool_stride + pr.coff;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = (wid & 1) * WM, wn = (wid >> 1) * WM;
    const int kt_n = (pr.k + BK - 1) / BK;
    const int64_t abase = mat * pack_tiles + pr.apk + (int64_t)(tl.tm / TM) * kt_n;
    const int64_t bbase = mat * pack_tiles + pr.bpk + (int64_t)(tl.tn / TM) * kt_n;
    const bf16x8_t *Ag = reinterpret_cast<const bf16x8_t *>(pack);
    acc4 acc[FM][FM];
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FM; j++) acc[i][j] = acc4{0, 0, 0, 0};
    bf16x8_t ra[CH], rb[CH];
    auto gload = [&](int kt) {
        const bf16x8_t *pa = Ag + (abase + kt) * (PK_TILE / 8), *pb = Ag + (bbase + kt) * (PK_TILE / 8);
#pragma unroll
        for (int h = 0; h < CH; h++) {
            ra[h] = pa[tid + 256 * h];
            rb[h] = pb[tid + 256 * h];
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int h = 0; h < CH; h++) {
            const int ch = tid + 256 * h, row = ch / (BK / 8), kc = (ch % (BK / 8)) * 8;
            *reinterpret_cast<bf16x8_t *>(&As[buf][row * LDK + kc]) = ra[h];
            *reinterpret_cast<bf16x8_t *>(&Bs[buf][row * LDK + kc]) = rb[h];
        }
    };
    gload(0);
    sstore(0);
    __syncthreads();
    const int srow = lane & 15, skq = (lane >> 4) * 8;
    for (int kt = 0; kt < kt_n; kt++) {
        const int cur = kt & 1;
        const float un = ldexpf(1.0f, pexp[abase + kt] + pexp[bbase + kt]);   // exact power of two
        if (kt + 1 < kt_n) gload(kt + 1);
        acc4 sl[FM][FM];
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
            for (int j = 0; j < FM; j++) sl[i][j] = acc4{0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < BK; ks += 32) {
            bf16x8_t av[FM], bv[FM];
#pragma unroll
            for (int i = 0; i < FM; i++)
                av[i] = *reinterpret_cast<const bf16x8_t *>(&As[cur][(wm + 16 * i + srow) * LDK + ks + skq]);
#pragma unroll
            for (int j = 0; j < FM; j++)
                bv[j] = *reinterpret_cast<const bf16x8_t *>(&Bs[cur][(wn + 16 * j + srow) * LDK + ks + skq]);
#pragma unroll
            for (int i = 0; i < FM; i++)
#pragma unroll
                for (int j = 0; j < FM; j++)
                    sl[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16v8_t, bv[j]),
                                                                       __builtin_bit_cast(bf16v8_t, av[i]),
                                                                       sl[i][j], 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
            for (int j = 0; j < FM; j++) acc[i][j] += sl[i][j] * un;
        if (kt + 1 < kt_n) sstore(cur ^ 1);
        __syncthreads();
    }
    // acc[i][j][q] = C[tm + wm + 16 i + lane % 16][tn + wn + 16 j + 4 (lane / 16) + q]
    const int gr0 = tl.tm + wm + srow;
#pragma unroll
    for (int j = 0; j < FM; j++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int gc = tl.tn + wn + 16 * j + 4 * (lane >> 4) + q;
            if (gc >= pr.n) continue;
            float *cp = C + (int64_t)gc * pr.ldc;
#pragma unroll
            for (int i = 0; i < FM; i++) {
                const int gr = gr0 + 16 * i;
                if (gr < pr.m) cp[gr] -= acc[i][j][q];
            }
        }
}

// ------------------------------------------------------------------------------------------
// End of a level: the factored panel [L11\U11; L21] and U12 of the large fronts leave the
// arena for the factor store (one workgroup per front column, the items of the level's column
// assembly: int2 {front, column}).  F22 stays in the arena as the contribution block.
// ------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_store(FactorArgs a)
{
    const int mat = blockIdx.y;
    if (skip_matrix(a.only, mat)) return;
    const int2 w = reinterpret_cast<const int2 *>(a.list)[blockIdx.x];
    const SnDev d = a.sn[w.x];
    const int c = w.y, nf = d.nf, np = d.np;
    const T *src = reinterpret_cast<const T *>(a.ws) + mat * a.ws_stride + d.woff + (int64_t)c * nf;
    T *store = reinterpret_cast<T *>(a.store) + mat * a.store_stride;
    if (c < np) {
        T *dst = store + d.foff + (int64_t)c * nf;
        for (int i = threadIdx.x; i < nf; i += blockDim.x) dst[i] = src[i];
    } else {
        T *dst = store + d.u12off + (int64_t)(c - np) * np;
        for (int i = threadIdx.x; i < np; i += blockDim.x) dst[i] = src[i];
    }
}

// Look-ahead: the first ndiag workgroups factor the diagonal block of block k of front
// list[blockIdx.x] (one wave; the other lanes leave at once), dispatched first; the rest apply
// the bulk of block k-1's trailing update (64 x 64 gemm tiles the next panel does not read).
// One LDS area serves either role.
template <typename T, int KB>
__global__ __launch_bounds__(256, sizeof(T) == 8 ? 2 : 4) void k_spec_diag_gemm(FactorArgs a, int k, float tau,
                                                        const GemmProb *__restrict__ probs,
                                                        const GemmTile *__restrict__ tiles, int ndiag, int dprob)
{
    __shared__ __align__(16) T lds[gemm_lds_elems<T, 64, 64>()];
    if (skip_matrix(a.only, blockIdx.y)) return;
    if ((int)blockIdx.x >= ndiag) {
        gemm_tile<T, 64, 64>(probs, tiles[blockIdx.x - ndiag], a.ws, a.ws_stride, lds);
        return;
    }
    if (dprob >= 0) {
        // this diagonal block's own share of the previous block's update, then its factorization
        gemm_tile<T, 64, 64>(probs, GemmTile{dprob + (int)blockIdx.x, 0, 0, 0}, a.ws, a.ws_stride, lds);
        __syncthreads();
    }
    static_assert(gemm_lds_elems<T, 64, 64>() >= KB * KB, "the multiplier slots fit in the tile's LDS");
    spec_diag_front<T, KB>(a, k, tau, (int)blockIdx.x, reinterpret_cast<typename SlotBits<T>::U *>(lds));
}

template <typename T>
hipError_t launch_gemm_tile(int tile, int use, const GemmProb *probs, const GemmTile *tiles, int count, void *pool,
                            int64_t pool_stride, int nmat, const int32_t *only, hipStream_t st, const void *pack,
                            const int32_t *pexp, int64_t pack_tiles)
{
    const dim3 grid(count, nmat), blk(256);
#define MF_GEMM_CASE(TI, TM, TN)                                                                                    \
    case TI:                                                                                                       \
        if (use == 2)                                                                                              \
            hipLaunchKernelGGL((k_gemm<T, TM, TN, 2>), grid, blk, 0, st, probs, tiles, pool, pool_stride, only);   \
        else if (use == 1)                                                                                         \
            hipLaunchKernelGGL((k_gemm<T, TM, TN, 1>), grid, blk, 0, st, probs, tiles, pool, pool_stride, only);   \
        else                                                                                                       \
            hipLaunchKernelGGL((k_gemm<T, TM, TN, 0>), grid, blk, 0, st, probs, tiles, pool, pool_stride, only);   \
        break;
    if (tile == GEMM_TILE_BF16) {
        if (sizeof(T) != 4) return hipErrorInvalidValue;
        if (!pack || !pexp) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_gemm_bf16, grid, blk, 0, st, probs, tiles, pool, pool_stride, only,
                           static_cast<const unsigned short *>(pack), pexp, pack_tiles);
        return hipGetLastError();
    }
    switch (tile) {
        MF_GEMM_CASE(0, 64, 64)
        MF_GEMM_CASE(1, 128, 64)
        MF_GEMM_CASE(2, 64, 128)
        MF_GEMM_CASE(3, 128, 128)
    default: return hipErrorInvalidValue;
    }
#undef MF_GEMM_CASE
    return hipGetLastError();
}
template hipError_t launch_gemm_tile<double>(int, int, const GemmProb *, const GemmTile *, int, void *, int64_t,
                                             int, const int32_t *, hipStream_t, const void *, const int32_t *, int64_t);
template hipError_t launch_gemm_tile<float>(int, int, const GemmProb *, const GemmTile *, int, void *, int64_t,
                                            int, const int32_t *, hipStream_t, const void *, const int32_t *, int64_t);

// ------------------------------------------------------------------------------------------
// Host launchers
// ------------------------------------------------------------------------------------------
template <typename T>
hipError_t launch_factor_step(const Launch &L, const FactorArgs &base, const int32_t *lists,
                              const GemmProb *probs, const GemmTile *tiles, int nmat,
                              hipStream_t st)
{
    FactorArgs a = base;
    a.list = lists + L.list_off;
    dim3 grid(L.count, nmat);
    switch (L.kind) {
    case L_FUSED_SMALL:
        hipLaunchKernelGGL(k_fused_small<T>, grid, dim3(256), L.shmem, st, a);
        break;
    case L_ASM_COLS:
        hipLaunchKernelGGL(k_asm_cols<T>, grid, dim3(256), L.shmem, st, a);
        break;
    case L_PANEL:
        switch (L.variant) {
        case 0: hipLaunchKernelGGL((k_panel<T, 256, 1>), grid, dim3(256), 0, st, a, L.k); break;
        case 1: hipLaunchKernelGGL((k_panel<T, 256, 2>), grid, dim3(256), 0, st, a, L.k); break;
        case 2: hipLaunchKernelGGL((k_panel<T, 256, 4>), grid, dim3(256), 0, st, a, L.k); break;
        case 3: hipLaunchKernelGGL((k_panel<T, 256, 8>), grid, dim3(256), 0, st, a, L.k); break;
        case 4: hipLaunchKernelGGL((k_panel<T, 512, 6>), grid, dim3(512), 0, st, a, L.k); break;
        case 5: hipLaunchKernelGGL((k_panel_big<T, 1024>), grid, dim3(1024), 0, st, a, L.k); break;
        default: return hipErrorInvalidValue;
        }
        break;
    case L_PANEL_APPLY:
        hipLaunchKernelGGL(k_panel_apply<T>, grid, dim3(256), 0, st, a, L.k);
        break;
    case L_INV_LEAF:
        hipLaunchKernelGGL(k_inv_leaf<T>, grid, dim3(128), 0, st, a);
        break;
    case L_SPEC_DIAG:
        hipLaunchKernelGGL((k_spec_diag<T, SPEC_KB>), grid, dim3(256), 0, st, a, L.k, L.tau);
        break;
    case L_SPEC_DIAG_GEMM:
        hipLaunchKernelGGL((k_spec_diag_gemm<T, SPEC_KB>), grid, dim3(256), 0, st, a, L.k, L.tau, probs + L.prob_off,
                           tiles + L.tile_off, L.count - L.ntiles, L.dprob);
        break;
    case L_SPEC_TRSM:
        hipLaunchKernelGGL((k_spec_trsm<T, SPEC_KB>), grid, dim3(4 * SPEC_ROWS), 0, st, a, L.k, L.tau,
                           reinterpret_cast<const T *>(a.spec));
        break;
    case L_GEMM:
        return launch_gemm_tile<T>(L.tile, L.variant == 3 ? 2 : L.variant == 1 ? 1 : 0, probs + L.prob_off,
                                   tiles + L.list_off, L.count, L.on_store ? base.store : base.ws,
                                   L.on_store ? base.store_stride : base.ws_stride, nmat, base.only, st, base.pack,
                                   base.pexp, base.pack_tiles);
    case L_PACK_BF16:
        if (sizeof(T) != 4) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_pack_bf16, grid, dim3(256), 0, st, a, probs + L.prob_off);
        break;
    case L_STORE:
        hipLaunchKernelGGL(k_store<T>, grid, dim3(256), 0, st, a);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template hipError_t launch_factor_step<double>(const Launch &, const FactorArgs &, const int32_t *,
                                               const GemmProb *, const GemmTile *, int, hipStream_t);
template hipError_t launch_factor_step<float>(const Launch &, const FactorArgs &, const int32_t *,
                                              const GemmProb *, const GemmTile *, int, hipStream_t);

// rows that one panel variant holds
int panel_variant_rows(int v)
{
    switch (v) {
    case 0: return 256;
    case 1: return 512;
    case 2: return 1024;
    case 3: return 2048;
    case 4: return 3072;
    case 5: return INT32_MAX;   // k_panel_big: any height
    default: return 0;
    }
}

}  // namespace mf
