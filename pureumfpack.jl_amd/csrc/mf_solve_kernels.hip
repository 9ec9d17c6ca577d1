// mf_solve_kernels.hip -- supernodal triangular solves, residual and refinement kernels.
//
// The solves read L and U straight from the front blocks.  Forward sweep (levels bottom-up,
// one workgroup per supernode): the front vector is the supernode's own entries of the
// permuted right-hand side plus the children's update vectors (the 1-D extend-add), the
// pivot part is permuted by the block pivoting, solved with the unit lower L11 and pushed
// through L21 into this supernode's update vector.  Backward sweep (levels top-down): each
// supernode gathers x at its update columns, subtracts U12 x_u and solves with U11.
// Together they compute what lsolve!/usolve! (src/solve.jl:10-39) compute column by column.
#include "mf_kernels.h"

namespace mf {

__global__ void k_fill_i32(int32_t *p, int32_t v, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

hipError_t launch_fill_i32(int32_t *p, int32_t v, int64_t n, hipStream_t st)
{
    int blocks = (int)std::min<int64_t>((n + 255) / 256, 1024);
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

template <typename T>
__global__ __launch_bounds__(256) void k_fwd(SolveArgs a)
{
    extern __shared__ unsigned char smem_raw[];
    const int mat = blockIdx.y;
    const SnDev d = a.sn[a.list[blockIdx.x]];
    const int nf = d.nf, np = d.np, nu = d.nu;
    T *fv = reinterpret_cast<T *>(smem_raw);
    T *ys = fv + nf;
    const T *F = reinterpret_cast<const T *>(a.pool) + mat * a.pool_stride + d.foff;
    const int32_t *prow = a.prow + mat * a.n_stride;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int64_t vb = (int64_t)mat * a.nrhs;
    for (int rhs = 0; rhs < a.nrhs; rhs++) {
        const T *w0 = reinterpret_cast<const T *>(a.w0) + (vb + rhs) * a.n;
        T *y = reinterpret_cast<T *>(a.y) + (vb + rhs) * a.n;
        T *uv = reinterpret_cast<T *>(a.uv) + (vb + rhs) * a.nupd;
        for (int i = tid; i < nf; i += nt) fv[i] = i < np ? w0[d.c1 + i] : T(0);
        __syncthreads();
        for (int ci = 0; ci < d.nchild; ci++) {
            const SnDev c = a.sn[a.child_list[d.child_off + ci]];
            const int32_t *rm = a.rmap + c.upd_off;
            for (int t = tid; t < c.nu; t += nt) fv[rm[t]] += uv[c.upd_off + t];
            __syncthreads();
        }
        for (int i = tid; i < np; i += nt) ys[i] = fv[prow[d.c1 + i] - d.c1];
        __syncthreads();
        for (int b0 = 0; b0 < np; b0 += 64) {
            const int bs = min(64, np - b0);
            if (tid < 64) {
                T v = tid < bs ? ys[b0 + tid] : T(0);
                for (int j = 0; j < bs; j++) {
                    T vj = __shfl(v, j);
                    if (tid > j && tid < bs) v -= F[(b0 + tid) + (int64_t)(b0 + j) * nf] * vj;
                }
                if (tid < bs) ys[b0 + tid] = v;
            }
            __syncthreads();
            for (int r = b0 + bs + tid; r < nf; r += nt) {
                T acc = T(0);
                for (int j = 0; j < bs; j++) acc += F[r + (int64_t)(b0 + j) * nf] * ys[b0 + j];
                if (r < np) ys[r] -= acc;
                else fv[r] -= acc;
            }
            __syncthreads();
        }
        for (int i = tid; i < np; i += nt) y[d.c1 + i] = ys[i];
        for (int t = tid; t < nu; t += nt) uv[d.upd_off + t] = fv[np + t];
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_bwd(SolveArgs a)
{
    extern __shared__ unsigned char smem_raw[];
    const int mat = blockIdx.y;
    const SnDev d = a.sn[a.list[blockIdx.x]];
    const int nf = d.nf, np = d.np, nu = d.nu;
    T *xs = reinterpret_cast<T *>(smem_raw);
    T *xu = xs + np;
    const T *F = reinterpret_cast<const T *>(a.pool) + mat * a.pool_stride + d.foff;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int64_t vb = (int64_t)mat * a.nrhs;
    for (int rhs = 0; rhs < a.nrhs; rhs++) {
        const T *y = reinterpret_cast<const T *>(a.y) + (vb + rhs) * a.n;
        T *x = reinterpret_cast<T *>(a.x) + (vb + rhs) * a.n;
        for (int t = tid; t < nu; t += nt) xu[t] = x[a.upd[d.upd_off + t]];
        __syncthreads();
        for (int i = tid; i < np; i += nt) {
            T acc = y[d.c1 + i];
            for (int t = 0; t < nu; t++) acc -= F[i + (int64_t)(np + t) * nf] * xu[t];
            xs[i] = acc;
        }
        __syncthreads();
        for (int b0 = ((np - 1) / 64) * 64; b0 >= 0; b0 -= 64) {
            const int bs = min(64, np - b0);
            if (tid < 64) {
                T v = tid < bs ? xs[b0 + tid] : T(0);
                for (int j = bs - 1; j >= 0; j--) {
                    if (tid == j) v = v / F[(b0 + j) + (int64_t)(b0 + j) * nf];
                    T vj = __shfl(v, j);
                    if (tid < j) v -= F[(b0 + tid) + (int64_t)(b0 + j) * nf] * vj;
                }
                if (tid < bs) xs[b0 + tid] = v;
            }
            __syncthreads();
            for (int r = tid; r < b0; r += nt) {
                T acc = T(0);
                for (int j = 0; j < bs; j++) acc += F[r + (int64_t)(b0 + j) * nf] * xs[b0 + j];
                xs[r] -= acc;
            }
            __syncthreads();
        }
        for (int i = tid; i < np; i += nt) x[d.c1 + i] = xs[i];
        __syncthreads();
    }
}

// w0[i] = Rs[qf[i]] * B[qf[i]] in the solve type: the scale-and-permute of
// src/interface.jl:128-130 expressed in V-row order.
template <typename T>
__global__ void k_prep_rhs(const double *__restrict__ B, int64_t ldb, const double *__restrict__ rs,
                           int64_t rs_stride, const int32_t *__restrict__ qf, T *__restrict__ w0,
                           int64_t n, int64_t nrhs, int64_t b_mat_stride)
{
    const int mat = blockIdx.y;
    const int64_t tot = n * nrhs;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = e % n, c = e / n;
        int g = qf[i];
        double v = B[mat * b_mat_stride + g + c * ldb];
        if (rs) v *= rs[mat * rs_stride + g];
        w0[(mat * nrhs + c) * n + i] = (T)v;
    }
}

// X[qf[k]] = x[k] (src/interface.jl:134-136).  accumulate=1 adds the correction in FP64
// (the refinement update x = x + d of src/interface.jl:174).
template <typename T>
__global__ void k_finish_x(const T *__restrict__ x, const int32_t *__restrict__ qf, double *__restrict__ X,
                           int64_t ldx, int64_t n, int64_t nrhs, int accumulate, int64_t x_mat_stride)
{
    const int mat = blockIdx.y;
    const int64_t tot = n * nrhs;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t k = e % n, c = e / n;
        double v = (double)x[(mat * nrhs + c) * n + k];
        double *dst = X + mat * x_mat_stride + qf[k] + c * ldx;
        *dst = accumulate ? *dst + v : v;
    }
}

// R = B - A X with A in CSR (row-parallel, fixed summation order per row).
__global__ void k_residual(const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                           const int32_t *__restrict__ vmap, const double *__restrict__ ax,
                           int64_t ax_stride, const double *__restrict__ X, int64_t ldx,
                           const double *__restrict__ B, int64_t ldb, double *__restrict__ R,
                           int64_t n, int64_t nrhs, int64_t mat_stride_x, int64_t mat_stride_b)
{
    const int mat = blockIdx.y;
    const double *axm = ax + mat * ax_stride;
    const int64_t tot = n * nrhs;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = e % n, c = e / n;
        const double *x = X + mat * mat_stride_x + c * ldx;
        double s = 0.0;
        for (int p = rp[i]; p < rp[i + 1]; p++) s += axm[vmap[p]] * x[ci[p]];
        R[mat * mat_stride_b + i + c * n] = B[mat * mat_stride_b + i + c * ldb] - s;
    }
}

// Per (matrix, rhs) column: out[0] = max|R|, out[1] = max|X|, out[2] = max|B|,
// out[3] = sum R^2, out[4] = sum B^2.  One workgroup per column.
__global__ __launch_bounds__(256) void k_norms(const double *__restrict__ R, const double *__restrict__ X,
                                               int64_t ldx, const double *__restrict__ B, int64_t ldb,
                                               int64_t n, int64_t nrhs, int64_t mat_stride_x,
                                               int64_t mat_stride_b, double *__restrict__ out)
{
    const int col = blockIdx.x, mat = blockIdx.y;
    const double *r = R + mat * mat_stride_b + (int64_t)col * n;
    const double *x = X + mat * mat_stride_x + (int64_t)col * ldx;
    const double *b = B + mat * mat_stride_b + (int64_t)col * ldb;
    double mr = 0, mx = 0, mb = 0, sr = 0, sb = 0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        double rv = r[i], xv = x[i], bv = b[i];
        mr = fmax(mr, fabs(rv)); mx = fmax(mx, fabs(xv)); mb = fmax(mb, fabs(bv));
        sr += rv * rv; sb += bv * bv;
    }
    __shared__ double sh[5][256];
    sh[0][threadIdx.x] = mr; sh[1][threadIdx.x] = mx; sh[2][threadIdx.x] = mb;
    sh[3][threadIdx.x] = sr; sh[4][threadIdx.x] = sb;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            for (int q = 0; q < 3; q++) sh[q][threadIdx.x] = fmax(sh[q][threadIdx.x], sh[q][threadIdx.x + s]);
            for (int q = 3; q < 5; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x < 5) out[((int64_t)mat * nrhs + col) * 5 + threadIdx.x] = sh[threadIdx.x][0];
}

// ||A||_inf per matrix (max row sum of |a_ij|), one workgroup per matrix.
__global__ __launch_bounds__(256) void k_anorm(const int32_t *__restrict__ rp, const int32_t *__restrict__ vmap,
                                               const double *__restrict__ ax, int64_t ax_stride, int64_t n,
                                               double *__restrict__ out)
{
    const double *axm = ax + blockIdx.x * ax_stride;
    double m = 0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        double s = 0;
        for (int p = rp[i]; p < rp[i + 1]; p++) s += fabs(axm[vmap[p]]);
        m = fmax(m, s);
    }
    __shared__ double sh[256];
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

// ---- CSC export (src/multifrontal.jl:277-350) ----
__global__ void k_rowfac(const int32_t *__restrict__ prow, int32_t *__restrict__ rowfac, int64_t n)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
        rowfac[prow[k]] = (int32_t)k;
}

// One workgroup per supernode: writes its L columns (diagonal first, pivot rows ascending,
// then its update rows relabelled through rowfac) and its U blocks at precomputed offsets.
template <typename T>
__global__ __launch_bounds__(256) void k_export(const SnDev *__restrict__ sn, const T *__restrict__ pool,
                                                const int32_t *__restrict__ upd, const int32_t *__restrict__ rowfac,
                                                const int64_t *__restrict__ Lp, const int64_t *__restrict__ Up,
                                                const int64_t *__restrict__ uoff, int64_t *__restrict__ Li,
                                                double *__restrict__ Lx, int64_t *__restrict__ Ui,
                                                double *__restrict__ Ux)
{
    const SnDev d = sn[blockIdx.x];
    const int nf = d.nf, np = d.np, nu = d.nu, c1 = d.c1;
    const T *F = pool + d.foff;
    for (int a = 0; a < np; a++) {
        const int g = c1 + a;
        const int64_t lp = Lp[g];
        const int cnt = 1 + (np - a - 1) + nu;
        for (int e = threadIdx.x; e < cnt; e += blockDim.x) {
            int64_t row;
            double v;
            if (e == 0) { row = g; v = 1.0; }
            else if (e < np - a) { row = g + e; v = (double)F[(a + e) + (int64_t)a * nf]; }
            else { int t = e - (np - a); row = rowfac[upd[d.upd_off + t]]; v = (double)F[(np + t) + (int64_t)a * nf]; }
            Li[lp + e] = row + 1;
            Lx[lp + e] = v;
        }
        const int64_t up = Up[g + 1] - (a + 1);
        for (int b = threadIdx.x; b <= a; b += blockDim.x) {
            Ui[up + b] = c1 + b + 1;
            Ux[up + b] = (double)F[b + (int64_t)a * nf];
        }
    }
    for (int t = 0; t < nu; t++) {
        const int64_t up = uoff[d.upd_off + t];
        for (int b = threadIdx.x; b < np; b += blockDim.x) {
            Ui[up + b] = c1 + b + 1;
            Ux[up + b] = (double)F[b + (int64_t)(np + t) * nf];
        }
    }
}

// ---- host launchers ----
template <typename T>
hipError_t launch_fwd(const SolveArgs &a, const int32_t *lists, int list_off, int count, int shmem, int nmat,
                      hipStream_t st)
{
    SolveArgs b = a;
    b.list = lists + list_off;
    hipLaunchKernelGGL(k_fwd<T>, dim3(count, nmat), dim3(256), shmem, st, b);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_bwd(const SolveArgs &a, const int32_t *lists, int list_off, int count, int shmem, int nmat,
                      hipStream_t st)
{
    SolveArgs b = a;
    b.list = lists + list_off;
    hipLaunchKernelGGL(k_bwd<T>, dim3(count, nmat), dim3(256), shmem, st, b);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_prep_rhs(const double *B, int64_t ldb, const double *rs, int64_t rs_stride, const int32_t *qf,
                           T *w0, int64_t n, int64_t nrhs, int64_t b_mat_stride, int nmat, hipStream_t st)
{
    int blocks = (int)std::min<int64_t>((n * nrhs + 255) / 256, 4096);
    hipLaunchKernelGGL(k_prep_rhs<T>, dim3(blocks, nmat), dim3(256), 0, st, B, ldb, rs, rs_stride, qf, w0, n, nrhs,
                       b_mat_stride);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_finish_x(const T *x, const int32_t *qf, double *X, int64_t ldx, int64_t n, int64_t nrhs,
                           int accumulate, int64_t x_mat_stride, int nmat, hipStream_t st)
{
    int blocks = (int)std::min<int64_t>((n * nrhs + 255) / 256, 4096);
    hipLaunchKernelGGL(k_finish_x<T>, dim3(blocks, nmat), dim3(256), 0, st, x, qf, X, ldx, n, nrhs, accumulate,
                       x_mat_stride);
    return hipGetLastError();
}
hipError_t launch_residual(const int32_t *rp, const int32_t *ci, const int32_t *vmap, const double *ax,
                           int64_t ax_stride, const double *X, int64_t ldx, const double *B, int64_t ldb, double *R,
                           int64_t n, int64_t nrhs, int64_t msx, int64_t msb, int nmat, hipStream_t st)
{
    int blocks = (int)std::min<int64_t>((n * nrhs + 255) / 256, 8192);
    hipLaunchKernelGGL(k_residual, dim3(blocks, nmat), dim3(256), 0, st, rp, ci, vmap, ax, ax_stride, X, ldx, B, ldb,
                       R, n, nrhs, msx, msb);
    return hipGetLastError();
}
hipError_t launch_norms(const double *R, const double *X, int64_t ldx, const double *B, int64_t ldb, int64_t n,
                        int64_t nrhs, int64_t msx, int64_t msb, double *out, int nmat, hipStream_t st)
{
    hipLaunchKernelGGL(k_norms, dim3((unsigned)nrhs, nmat), dim3(256), 0, st, R, X, ldx, B, ldb, n, nrhs, msx, msb,
                       out);
    return hipGetLastError();
}
hipError_t launch_anorm(const int32_t *rp, const int32_t *vmap, const double *ax, int64_t ax_stride, int64_t n,
                        double *out, int nmat, hipStream_t st)
{
    hipLaunchKernelGGL(k_anorm, dim3(nmat), dim3(256), 0, st, rp, vmap, ax, ax_stride, n, out);
    return hipGetLastError();
}
hipError_t launch_rowfac(const int32_t *prow, int32_t *rowfac, int64_t n, hipStream_t st)
{
    int blocks = (int)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_rowfac, dim3(blocks), dim3(256), 0, st, prow, rowfac, n);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_export(const SnDev *sn, int64_t nsuper, const T *pool, const int32_t *upd, const int32_t *rowfac,
                         const int64_t *Lp, const int64_t *Up, const int64_t *uoff, int64_t *Li, double *Lx,
                         int64_t *Ui, double *Ux, hipStream_t st)
{
    hipLaunchKernelGGL(k_export<T>, dim3((unsigned)nsuper), dim3(256), 0, st, sn, pool, upd, rowfac, Lp, Up, uoff, Li,
                       Lx, Ui, Ux);
    return hipGetLastError();
}

#define MF_INST(T)                                                                                                \
    template hipError_t launch_fwd<T>(const SolveArgs &, const int32_t *, int, int, int, int, hipStream_t);     \
    template hipError_t launch_bwd<T>(const SolveArgs &, const int32_t *, int, int, int, int, hipStream_t);     \
    template hipError_t launch_prep_rhs<T>(const double *, int64_t, const double *, int64_t, const int32_t *,   \
                                           T *, int64_t, int64_t, int64_t, int, hipStream_t);                   \
    template hipError_t launch_finish_x<T>(const T *, const int32_t *, double *, int64_t, int64_t, int64_t,     \
                                           int, int64_t, int, hipStream_t);                                     \
    template hipError_t launch_export<T>(const SnDev *, int64_t, const T *, const int32_t *, const int32_t *,   \
                                         const int64_t *, const int64_t *, const int64_t *, int64_t *, double *, \
                                         int64_t *, double *, hipStream_t);
MF_INST(double)
MF_INST(float)
#undef MF_INST

}  // namespace mf
