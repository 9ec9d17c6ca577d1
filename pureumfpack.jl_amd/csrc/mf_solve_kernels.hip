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

__global__ void k_fill_i32_masked(int32_t *p, int32_t v, const int32_t *mask, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (!mask || mask[i]) p[i] = v;
}

hipError_t launch_fill_i32_masked(int32_t *p, int32_t v, const int32_t *mask, int64_t n, hipStream_t st)
{
    int blocks = (int)std::min<int64_t>((n + 255) / 256, 1024);
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_fill_i32_masked, dim3(blocks), dim3(256), 0, st, p, v, mask, n);
    return hipGetLastError();
}

hipError_t launch_fill_i32(int32_t *p, int32_t v, int64_t n, hipStream_t st)
{
    int blocks = (int)std::min<int64_t>((n + 255) / 256, 1024);
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

constexpr int SOLVE_RC = 8;   // right-hand sides processed together per pass
constexpr int RW = SOLVE_RC / 4;   // right-hand sides per wave in the diagonal-block solves
constexpr int SOLVE_NB = 64;  // triangular block size (one wave solves a diagonal block)

// One wave: v <- L^-1 v (R right-hand sides) for the unit lower bs x bs block staged in sd
// (leading dimension ldd, lane = row; rows and columns >= bs of sd and v are zero).  Columns in groups of 8,
// read from LDS before their 8 steps, without branches: a step past bs has a zero column and
// changes nothing.  The value of lane j reaches the other lanes by v_readlane (j is
// wave-uniform).
template <typename T, int R>
__device__ __forceinline__ void wave_lower_solve(const T *sd, int ldd, int bs, int lane, T (&v)[R])
{
    for (int j0 = 0; j0 < bs; j0 += 8) {
        T l[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const T t = sd[lane + (j0 + q) * ldd];
            l[q] = lane > j0 + q ? t : T(0);
        }
#pragma unroll
        for (int q = 0; q < 8; q++)
#pragma unroll
            for (int r = 0; r < R; r++) v[r] -= l[q] * lane_value(v[r], j0 + q);
    }
}

// One wave: v <- U^-1 v for the upper bs x bs block staged in sd, last column first; lane j
// takes x_j = v_j / u_jj (one division per step), the lanes above subtract u_ij x_j.
// Each lane divides once (its own 1/u_jj); step j takes lane j's by v_readlane.  Steps below
// column 0 (the last group) have u = 0 and 1/u_jj = 0 and change nothing.
template <typename T, int R>
__device__ __forceinline__ void wave_upper_solve(const T *sd, int ldd, int bs, int lane, T (&v)[R])
{
    const T rd = lane < bs ? T(1) / sd[lane + lane * ldd] : T(1);
    for (int j1 = bs - 1; j1 >= 0; j1 -= 8) {
        T u[8], rj[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int j = max(j1 - q, 0);
            const T t = sd[lane + j * ldd];
            u[q] = lane < j1 - q ? t : T(0);
            rj[q] = j1 - q >= 0 ? lane_value(rd, j) : T(0);
        }
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int j = j1 - q;
#pragma unroll
            for (int r = 0; r < R; r++) {
                const T t = lane_value(v[r], max(j, 0)) * rj[q];
                v[r] = lane == j ? t : v[r] - u[q] * t;
            }
        }
    }
}

// Forward sweep, part 0: pull the children's update vectors into this supernode's front
// vector, one thread per front position (items int4 {front, pos0}): pivot rows accumulate in
// w0 (V-row order), update rows start this supernode's update vector uv.  The children of a
// position come from the extend-add lists of the factorization, in the same reverse child
// order, so the sums do not depend on scheduling.
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_gather(SolveArgs a, const int32_t *__restrict__ ecolptr,
                                                    const int32_t *__restrict__ eupd)
{
    const int mat = blockIdx.y;
    const int4 w = reinterpret_cast<const int4 *>(a.list)[blockIdx.x];
    const SnDev d = a.sn[w.x];
    const int p = w.y + threadIdx.x;
    if (p >= d.nf) return;
    const int64_t vb = (int64_t)mat * a.nrhs;
    T *w0 = reinterpret_cast<T *>(a.w0) + vb * a.n;
    T *uv = reinterpret_cast<T *>(a.uv) + vb * a.nupd;
    const int q0 = ecolptr[d.fcol + p], q1 = ecolptr[d.fcol + p + 1];
    // the offsets of the first GQ contributions, then all of their values, are requested together
    constexpr int GQ = 4;
    int32_t offs[GQ];
#pragma unroll
    for (int u = 0; u < GQ; u++) offs[u] = q0 + u < q1 ? eupd[q0 + u] : 0;
    for (int r0 = blockIdx.z * SOLVE_RC; r0 < (int)a.nrhs; r0 += gridDim.z * SOLVE_RC) {
        const int rc = min(SOLVE_RC, (int)a.nrhs - r0);
        T acc[SOLVE_RC];
#pragma unroll
        for (int r = 0; r < SOLVE_RC; r++)
            acc[r] = (r < rc && p < d.np) ? w0[(int64_t)(r0 + r) * a.n + d.c1 + p] : T(0);
        T v[GQ][SOLVE_RC];
#pragma unroll
        for (int u = 0; u < GQ; u++)
#pragma unroll
            for (int r = 0; r < SOLVE_RC; r++)
                v[u][r] = (q0 + u < q1 && r < rc) ? uv[(int64_t)(r0 + r) * a.nupd + offs[u]] : T(0);
#pragma unroll
        for (int u = 0; u < GQ; u++)
            if (q0 + u < q1) {
#pragma unroll
                for (int r = 0; r < SOLVE_RC; r++) acc[r] += v[u][r];
            }
        for (int q = q0 + GQ; q < q1; q++) {
            const int32_t off = eupd[q];
#pragma unroll
            for (int r = 0; r < SOLVE_RC; r++)
                if (r < rc) acc[r] += uv[(int64_t)(r0 + r) * a.nupd + off];
        }
#pragma unroll
        for (int r = 0; r < SOLVE_RC; r++) {
            if (r >= rc) continue;
            if (p < d.np) w0[(int64_t)(r0 + r) * a.n + d.c1 + p] = acc[r];
            else uv[(int64_t)(r0 + r) * a.nupd + d.upd_off + p - d.np] = acc[r];
        }
    }
}

// Forward sweep, part 1, fronts without stored inverses (np <= INV_NP, two diagonal blocks at
// most): one workgroup per front permutes the gathered pivot part by the block pivoting into
// factor order (in LDS), solves with the unit lower L11 and writes y.  The L21 product that
// finishes the update vector is k_rect_part kind 0.  Fronts with inverses skip this:
// k_rect_part kind 2 reads the gathered vector through the pivoting itself.  Every factor
// value the workgroup needs is loaded before the vector arrives.
constexpr int TRI_NP = INV_NP;
static_assert(TRI_NP == 2 * SOLVE_NB, "the triangle kernels handle two diagonal blocks");

template <typename T>
__global__ __launch_bounds__(256) void k_fwd_tri(SolveArgs a)
{
    constexpr int NT = 256, LDD = SOLVE_NB + 1, DPT = SOLVE_NB * SOLVE_NB / NT;
    __shared__ T g[SOLVE_RC * TRI_NP];
    __shared__ T sd[SOLVE_NB * LDD];
    __shared__ T red[SOLVE_RC * NT];
    const int mat = blockIdx.y;
    const SnDev d = a.sn[a.list[blockIdx.x]];
    if (d.ioff >= 0) return;   // large front: inv(L11) is applied by k_rect_part kind 2
    const int nf = d.nf, np = d.np, c1 = d.c1;
    const T *F = reinterpret_cast<const T *>(a.store) + mat * a.store_stride + d.foff;
    const int32_t *prow = a.prow + mat * a.n_stride;
    const int tid = threadIdx.x, lane = tid & 63;
    const int64_t vb = (int64_t)mat * a.nrhs;
    const T *w0 = reinterpret_cast<const T *>(a.w0) + vb * a.n;
    T *y = reinterpret_cast<T *>(a.y) + vb * a.n;
    const int nrhs = (int)a.nrhs;
    const int bs0 = min(SOLVE_NB, np), bs1 = np - bs0;   // bs1 > 0: a second block
    // diagonal blocks of L11 and the block L11[64 + ur, 16 uq .. 16 uq + 15] below the first
    T d0[DPT], d1[DPT], lr[16];
    const int ur = tid & 63, uq = tid >> 6;
#pragma unroll
    for (int h = 0; h < DPT; h++) {
        const int e = tid + h * NT, i = e % SOLVE_NB, j = e / SOLVE_NB;
        d0[h] = (i < bs0 && j < bs0) ? F[i + (int64_t)j * nf] : T(0);
        d1[h] = (i < bs1 && j < bs1) ? F[(SOLVE_NB + i) + (int64_t)(SOLVE_NB + j) * nf] : T(0);
    }
#pragma unroll
    for (int q = 0; q < 16; q++)
        lr[q] = ur < bs1 ? F[(SOLVE_NB + ur) + (int64_t)(uq * 16 + q) * nf] : T(0);
    const int64_t src = tid < np ? prow[c1 + tid] : 0;
    for (int r0 = blockIdx.z * SOLVE_RC; r0 < nrhs; r0 += gridDim.z * SOLVE_RC) {
        const int rc = min(SOLVE_RC, nrhs - r0);
        if (tid < TRI_NP) {
#pragma unroll
            for (int r = 0; r < SOLVE_RC; r++)
                g[r * TRI_NP + tid] = (r < rc && tid < np) ? w0[(int64_t)(r0 + r) * a.n + src] : T(0);
        }
#pragma unroll
        for (int h = 0; h < DPT; h++) {
            const int e = tid + h * NT;
            sd[(e % SOLVE_NB) + (e / SOLVE_NB) * LDD] = d0[h];
        }
        __syncthreads();
        {
            // the right-hand sides split over the four waves, RW each
            const int rw0 = (tid >> 6) * RW;
            T v[RW];
#pragma unroll
            for (int rr = 0; rr < RW; rr++) {
                const int r = rw0 + rr;
                v[rr] = lane < bs0 ? g[r * TRI_NP + lane] : T(0);
            }
            wave_lower_solve<T, RW>(sd, LDD, bs0, lane, v);
            if (lane < bs0) {
#pragma unroll
                for (int rr = 0; rr < RW; rr++) {
                    const int r = rw0 + rr;
                    g[r * TRI_NP + lane] = v[rr];
                }
            }
        }
        __syncthreads();
        if (bs1 > 0) {
            // rows 64.. -= L11[64.., 0:64] x_0, four partial sums per row added in order
            T acc[SOLVE_RC] = {};
#pragma unroll
            for (int q = 0; q < 16; q++)
#pragma unroll
                for (int r = 0; r < SOLVE_RC; r++) acc[r] += lr[q] * g[r * TRI_NP + uq * 16 + q];
#pragma unroll
            for (int r = 0; r < SOLVE_RC; r++) red[r * NT + tid] = acc[r];
#pragma unroll
            for (int h = 0; h < DPT; h++) {
                const int e = tid + h * NT;
                sd[(e % SOLVE_NB) + (e / SOLVE_NB) * LDD] = d1[h];
            }
            __syncthreads();
            {
                const int rw0 = (tid >> 6) * RW;
                T v[RW];
#pragma unroll
                for (int rr = 0; rr < RW; rr++) {
                    const int r = rw0 + rr;
                    const T sum = ((red[r * NT + lane] + red[r * NT + 64 + lane]) + red[r * NT + 128 + lane]) +
                                  red[r * NT + 192 + lane];
                    v[rr] = lane < bs1 ? g[r * TRI_NP + SOLVE_NB + lane] - sum : T(0);
                }
                wave_lower_solve<T, RW>(sd, LDD, bs1, lane, v);
                if (lane < bs1) {
#pragma unroll
                    for (int rr = 0; rr < RW; rr++) g[(rw0 + rr) * TRI_NP + SOLVE_NB + lane] = v[rr];
                }
            }
            __syncthreads();
        }
        if (tid < np) {
#pragma unroll
            for (int r = 0; r < SOLVE_RC; r++)
                if (r < rc) y[(int64_t)(r0 + r) * a.n + c1 + tid] = g[r * TRI_NP + tid];
        }
        __syncthreads();
    }
}

// Rectangular products of the sweeps, split over (256-row block) x (SOLVE_CK-column chunk)
// workgroups:
//   kind 0: uv[t] -= sum_a L21[t, a] y[c1 + a]                  rows t < nu, columns a < np
//   kind 1: x[c1+i] = y[c1+i] - sum_t U12[i, t] x[upd[t]]      rows i < np, columns t < nu
//           (fronts with inverses: into zbuf, inv(U11) follows)
//   kind 2: y[c1+i] = sum_a inv(L11)[i, a] w0[prow[c1 + a]]     lower-triangle chunks only
//   kind 3: x[c1+i] = sum_a inv(U11)[i, a] zbuf[c1 + a]         upper-triangle chunks only
// k_rect_part items are int4 {front, row0, col0, slot}: two workgroups per item (128 rows
// each) write its partial sums (256 rows x nrhs) to partial slot `slot`.  k_rect_sum items are {dst, rows | zb << 16,
// nchunks, slot0}: the row block's chunks own slots slot0 .. slot0 + nchunks - 1, added in a
// fixed order (sixteen interleaved partial sums, each in chunk order, then the sixteen in
// order: the result does not depend on scheduling) and applied to `rows` entries from dst
// (c1 + row0, or upd_off + row0 for kind 0; zb: kind 1 writes zbuf).
constexpr int SOLVE_CK = 64;
constexpr int RECT_SUM_TPR = 16;   // threads per row of a split-K sum (and its workgroups per item)

// acc[r] += sum_{j < bs} M[j * ld] * s[r * sld + j], G loads in flight
template <typename T, int G>
__device__ __forceinline__ void dot_rows(const T *M, int64_t ld, int bs, const T *s, int sld, T (&acc)[SOLVE_RC])
{
    int j = 0;
    for (; j + G <= bs; j += G) {
        T l[G];
#pragma unroll
        for (int q = 0; q < G; q++) l[q] = M[(int64_t)(j + q) * ld];
#pragma unroll
        for (int q = 0; q < G; q++)
#pragma unroll
            for (int r = 0; r < SOLVE_RC; r++) acc[r] += l[q] * s[r * sld + j + q];
    }
    for (; j < bs; j++) {
        const T l = M[(int64_t)j * ld];
#pragma unroll
        for (int r = 0; r < SOLVE_RC; r++) acc[r] += l * s[r * sld + j];
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_rect_part(SolveArgs a, int kind)
{
    // two workgroups per item (128 rows each); two threads per row, each with half of the chunk's
    // columns, loaded before the vector is staged; the two halves added in order in LDS
    constexpr int HC = SOLVE_CK / 2;
    __shared__ T sv[SOLVE_CK * SOLVE_RC];
    __shared__ T red[SOLVE_RC * 128];
    const int mat = blockIdx.y;
    const int4 w = reinterpret_cast<const int4 *>(a.list)[blockIdx.x >> 1];
    const SnDev d = a.sn[w.x];
    const int nf = d.nf, np = d.np, nu = d.nu, c1 = d.c1;
    const T *F = reinterpret_cast<const T *>(a.store) + mat * a.store_stride + d.foff;
    const int tid = threadIdx.x, lr = tid & 127, hf = tid >> 7;
    const int ir = ((int)blockIdx.x & 1) * 128 + lr;   // row within the item's 256
    const int row = w.y + ir;
    const int nrows = kind == 0 ? nu : np, ncols = kind == 1 ? nu : np;
    const int col0 = w.z;
    const int bs = max(0, min(SOLVE_CK, ncols - col0));
    const int hb = max(0, min(HC, bs - hf * HC));   // this thread's columns
    const int64_t vb = (int64_t)mat * a.nrhs;
    const T *y = reinterpret_cast<const T *>(a.y) + vb * a.n;
    const T *x = reinterpret_cast<const T *>(a.x) + vb * a.n;
    const T *zb = reinterpret_cast<const T *>(a.zbuf) + vb * a.n;
    const T *w0 = reinterpret_cast<const T *>(a.w0) + vb * a.n;
    const int32_t *prow = a.prow + mat * a.n_stride;
    T *part = reinterpret_cast<T *>(a.part) + mat * a.part_stride + (int64_t)w.w * 256 * a.nrhs;
    const int32_t *upd = a.upd + d.upd_off;
    // row r of the operand: kind 0 L21 row (np + r) over columns a (ld nf); kind 1 U12 row r
    // over columns t (ld np); kinds 2 / 3 row r of inv(L11) / inv(U11) (ld np)
    const T *store = reinterpret_cast<const T *>(a.store) + mat * a.store_stride;
    const T *rowp;
    int ld = np;
    switch (kind) {
    case 0: rowp = F + (np + row); ld = nf; break;
    case 1: rowp = store + d.u12off + row; break;
    case 2: rowp = store + d.ioff + row; break;
    default: rowp = store + d.ioff + (int64_t)np * np + row; break;
    }
    const int nrhs = (int)a.nrhs;
    T l[HC];
    {
        const T *mrow = rowp + (int64_t)(col0 + hf * HC) * ld;
        const bool on = row < nrows;
#pragma unroll
        for (int q = 0; q < HC; q++) l[q] = (on && q < hb) ? mrow[(int64_t)q * ld] : T(0);
    }
    for (int r0 = blockIdx.z * SOLVE_RC; r0 < nrhs; r0 += gridDim.z * SOLVE_RC) {
        const int rc = min(SOLVE_RC, nrhs - r0);
        __syncthreads();
#pragma unroll
        for (int h = 0; h < SOLVE_CK * SOLVE_RC / 256; h++) {
            const int e = tid + h * 256;
            const int j = e % SOLVE_CK, r = e / SOLVE_CK;
            T v = T(0);
            if (j < bs && r < rc) {
                const int64_t ro = (int64_t)(r0 + r) * a.n;
                v = kind == 0   ? y[ro + c1 + col0 + j]
                    : kind == 1 ? x[ro + upd[col0 + j]]
                    : kind == 2 ? w0[ro + prow[c1 + col0 + j]]
                                : zb[ro + c1 + col0 + j];
            }
            sv[e] = v;
        }
        __syncthreads();
        T acc[SOLVE_RC] = {};
#pragma unroll
        for (int q = 0; q < HC; q++)
#pragma unroll
            for (int r = 0; r < SOLVE_RC; r++) acc[r] += l[q] * sv[r * SOLVE_CK + hf * HC + q];
        if (hf == 1) {
#pragma unroll
            for (int r = 0; r < SOLVE_RC; r++) red[r * 128 + lr] = acc[r];
        }
        __syncthreads();
        if (hf == 0) {
#pragma unroll
            for (int r = 0; r < SOLVE_RC; r++)
                if (r < rc) part[(int64_t)(r0 + r) * 256 + ir] = acc[r] + red[r * 128 + lr];
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_rect_sum(SolveArgs a, int kind)
{
    // sixteen workgroups per item (16 rows each); sixteen threads per row, thread g adding the
    // chunks c = g, g + 16, g + 32, ... in order, all of a thread's loads in flight together for
    // up to 64 chunks; the sixteen partial sums of a row are added in order in LDS
    constexpr int TPR = RECT_SUM_TPR, RPW = 256 / TPR, U = 4;
    __shared__ T red[TPR * SOLVE_RC * RPW];
    const int mat = blockIdx.y;
    const int4 w = reinterpret_cast<const int4 *>(a.list)[blockIdx.x / TPR];
    const int lr = threadIdx.x % RPW, g = threadIdx.x / RPW;
    const int rr = ((int)blockIdx.x % TPR) * RPW + lr;
    const bool on = rr < (w.y & 0xffff);
    const int64_t dst = (int64_t)w.x + rr;
    const bool to_zb = (w.y >> 16) != 0;
    const int nch = w.z;
    const int64_t vb = (int64_t)mat * a.nrhs;
    T *y = reinterpret_cast<T *>(a.y) + vb * a.n;
    T *x = reinterpret_cast<T *>(a.x) + vb * a.n;
    T *zb = reinterpret_cast<T *>(a.zbuf) + vb * a.n;
    T *uv = reinterpret_cast<T *>(a.uv) + vb * a.nupd;
    const int64_t cs = (int64_t)256 * a.nrhs;   // one partial slot
    const T *part = reinterpret_cast<const T *>(a.part) + mat * a.part_stride + (int64_t)w.w * cs;
    const int nrhs = (int)a.nrhs;
    for (int r0 = blockIdx.z * SOLVE_RC; r0 < nrhs; r0 += gridDim.z * SOLVE_RC) {
        const int rc = min(SOLVE_RC, nrhs - r0);
        const T *pp = part + (int64_t)r0 * 256 + rr;
        // the values the sums are applied to (kinds 0 and 1) do not depend on the partial sums:
        // they are read first
        T base[SOLVE_RC];
#pragma unroll
        for (int r = 0; r < SOLVE_RC; r++) {
            base[r] = T(0);
            if (g == 0 && on && r < rc) {
                if (kind == 0) base[r] = uv[(int64_t)(r0 + r) * a.nupd + dst];
                else if (kind == 1) base[r] = y[(int64_t)(r0 + r) * a.n + dst];
            }
        }
        T s[SOLVE_RC] = {};
        if (on) {
            for (int c = g; c < nch; c += U * TPR) {
                T v[U][SOLVE_RC];
#pragma unroll
                for (int q = 0; q < U; q++)
#pragma unroll
                    for (int r = 0; r < SOLVE_RC; r++)
                        v[q][r] = (c + q * TPR < nch && r < rc) ? pp[(int64_t)(c + q * TPR) * cs + r * 256] : T(0);
#pragma unroll
                for (int q = 0; q < U; q++)
#pragma unroll
                    for (int r = 0; r < SOLVE_RC; r++) s[r] += v[q][r];
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < SOLVE_RC; r++) red[(g * SOLVE_RC + r) * RPW + lr] = s[r];
        __syncthreads();
        if (g == 0 && on) {
#pragma unroll
            for (int r = 0; r < SOLVE_RC; r++) {
                if (r >= rc) continue;
                T sum = red[r * RPW + lr];
#pragma unroll
                for (int q = 1; q < TPR; q++) sum += red[(q * SOLVE_RC + r) * RPW + lr];
                const int64_t ro = (int64_t)(r0 + r) * a.n + dst;
                switch (kind) {
                case 0: uv[(int64_t)(r0 + r) * a.nupd + dst] = base[r] - sum; break;
                case 1: (to_zb ? zb : x)[ro] = base[r] - sum; break;   // inv(U11) next
                case 2: y[ro] = sum; break;
                default: x[ro] = sum; break;
                }
            }
        }
    }
}

// Forward L21 product without split-K, for levels whose split fronts all have np <= 4 *
// SOLVE_CK: uv[t] -= sum_a L21[t, a] y[c1 + a].  Items int4 {front, row0} of 64 rows.
template <typename T>
__global__ __launch_bounds__(256) void k_rect_direct(SolveArgs a)
{
    // two workgroups per item (32 rows each); eight column groups of 32 per row, each thread's
    // operand values loaded before the vector is staged; the eight partial sums of a row are
    // added in group order in LDS
    constexpr int NG = 8, CW = 32, NC = NG * CW;
    static_assert(NC == 4 * SOLVE_CK, "the direct product covers pivot blocks of four chunks");
    __shared__ T sv[NC * SOLVE_RC];   // y, columns 0 .. 255, by right-hand side
    __shared__ T red[NG * SOLVE_RC * 32];
    const int mat = blockIdx.y;
    const int4 w = reinterpret_cast<const int4 *>(a.list)[blockIdx.x >> 1];
    const SnDev d = a.sn[w.x];
    const int nf = d.nf, np = d.np, nu = d.nu, c1 = d.c1;
    const T *F = reinterpret_cast<const T *>(a.store) + mat * a.store_stride + d.foff;
    const int tid = threadIdx.x, lr = tid & 31, cg = tid >> 5;
    const int row = w.y + ((int)blockIdx.x & 1) * 32 + lr;
    const int64_t vb = (int64_t)mat * a.nrhs;
    const T *y = reinterpret_cast<const T *>(a.y) + vb * a.n;
    T *uv = reinterpret_cast<T *>(a.uv) + vb * a.nupd;
    const int col0 = cg * CW;
    const int bs = max(0, min(CW, np - col0));
    T l[CW];
    {
        const T *mrow = F + (np + row) + (int64_t)col0 * nf;
        const bool on = row < nu;
#pragma unroll
        for (int q = 0; q < CW; q++) l[q] = (on && q < bs) ? mrow[(int64_t)q * nf] : T(0);
    }
    const int nrhs = (int)a.nrhs;
    for (int r0 = blockIdx.z * SOLVE_RC; r0 < nrhs; r0 += gridDim.z * SOLVE_RC) {
        const int rc = min(SOLVE_RC, nrhs - r0);
        __syncthreads();
#pragma unroll
        for (int h = 0; h < NC * SOLVE_RC / 256; h++) {
            const int e = tid + h * 256;
            const int j = e % NC, r = e / NC;
            sv[e] = (j < np && r < rc) ? y[(int64_t)(r0 + r) * a.n + c1 + j] : T(0);
        }
        __syncthreads();
        T acc[SOLVE_RC] = {};
#pragma unroll
        for (int q = 0; q < CW; q++)
#pragma unroll
            for (int r = 0; r < SOLVE_RC; r++) acc[r] += l[q] * sv[r * NC + col0 + q];
#pragma unroll
        for (int r = 0; r < SOLVE_RC; r++) red[(cg * SOLVE_RC + r) * 32 + lr] = acc[r];
        __syncthreads();
        if (cg == 0 && row < nu) {
#pragma unroll
            for (int r = 0; r < SOLVE_RC; r++) {
                if (r >= rc) continue;
                T sum = red[r * 32 + lr];
#pragma unroll
                for (int q = 1; q < NG; q++) sum += red[(q * SOLVE_RC + r) * 32 + lr];
                uv[(int64_t)(r0 + r) * a.nupd + d.upd_off + row] -= sum;
            }
        }
    }
}

// Backward sweep, part 2, fronts without stored inverses (np <= TRI_NP): x_s <- U11^-1 x_s,
// bottom diagonal block first, in LDS; the factor values are loaded before the vector.
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_tri(SolveArgs a)
{
    constexpr int NT = 256, LDD = SOLVE_NB + 1, DPT = SOLVE_NB * SOLVE_NB / NT;
    __shared__ T g[SOLVE_RC * TRI_NP];
    __shared__ T sd[SOLVE_NB * LDD];
    __shared__ T red[SOLVE_RC * NT];
    const int mat = blockIdx.y;
    const SnDev d = a.sn[a.list[blockIdx.x]];
    if (d.ioff >= 0) return;  // large front: inv(U11) was applied by k_rect_part kind 3
    const int nf = d.nf, np = d.np, c1 = d.c1;
    const T *F = reinterpret_cast<const T *>(a.store) + mat * a.store_stride + d.foff;
    const int tid = threadIdx.x, lane = tid & 63;
    const int64_t vb = (int64_t)mat * a.nrhs;
    T *x = reinterpret_cast<T *>(a.x) + vb * a.n;
    const int nrhs = (int)a.nrhs;
    const int bs0 = min(SOLVE_NB, np), bs1 = np - bs0;
    // diagonal blocks of U11 and U11[ur, 64 + 16 uq .. 64 + 16 uq + 15] above the second
    T d0[DPT], d1[DPT], ur16[16];
    const int ur = tid & 63, uq = tid >> 6;
#pragma unroll
    for (int h = 0; h < DPT; h++) {
        const int e = tid + h * NT, i = e % SOLVE_NB, j = e / SOLVE_NB;
        d0[h] = (i < bs0 && j < bs0) ? F[i + (int64_t)j * nf] : T(0);
        d1[h] = (i < bs1 && j < bs1) ? F[(SOLVE_NB + i) + (int64_t)(SOLVE_NB + j) * nf] : T(0);
    }
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int c = uq * 16 + q;
        ur16[q] = (ur < bs0 && c < bs1) ? F[ur + (int64_t)(SOLVE_NB + c) * nf] : T(0);
    }
    for (int r0 = blockIdx.z * SOLVE_RC; r0 < nrhs; r0 += gridDim.z * SOLVE_RC) {
        const int rc = min(SOLVE_RC, nrhs - r0);
        // every row of g is written (zero past np): the U11 coupling product below reads the
        // whole second block, with zero factor entries past np, and 0 * (stale LDS) need
        // not be 0
        if (tid < TRI_NP) {
#pragma unroll
            for (int r = 0; r < SOLVE_RC; r++)
                g[r * TRI_NP + tid] = (r < rc && tid < np) ? x[(int64_t)(r0 + r) * a.n + c1 + tid] : T(0);
        }
        if (bs1 > 0) {
#pragma unroll
            for (int h = 0; h < DPT; h++) {
                const int e = tid + h * NT;
                sd[(e % SOLVE_NB) + (e / SOLVE_NB) * LDD] = d1[h];
            }
            __syncthreads();
            {
                // the right-hand sides split over the four waves, RW each
                const int rw0 = (tid >> 6) * RW;
                T v[RW];
#pragma unroll
                for (int rr = 0; rr < RW; rr++) {
                    const int r = rw0 + rr;
                    v[rr] = lane < bs1 ? g[r * TRI_NP + SOLVE_NB + lane] : T(0);
                }
                wave_upper_solve<T, RW>(sd, LDD, bs1, lane, v);
                if (lane < bs1) {
#pragma unroll
                    for (int rr = 0; rr < RW; rr++) {
                        const int r = rw0 + rr;
                        g[r * TRI_NP + SOLVE_NB + lane] = v[rr];
                    }
                }
            }
            __syncthreads();
            // rows 0..63 -= U11[0:64, 64..] x_1, four partial sums per row added in order
            T acc[SOLVE_RC] = {};
#pragma unroll
            for (int q = 0; q < 16; q++)
#pragma unroll
                for (int r = 0; r < SOLVE_RC; r++) acc[r] += ur16[q] * g[r * TRI_NP + SOLVE_NB + uq * 16 + q];
#pragma unroll
            for (int r = 0; r < SOLVE_RC; r++) red[r * NT + tid] = acc[r];
        }
#pragma unroll
        for (int h = 0; h < DPT; h++) {
            const int e = tid + h * NT;
            sd[(e % SOLVE_NB) + (e / SOLVE_NB) * LDD] = d0[h];
        }
        __syncthreads();
        {
            const int rw0 = (tid >> 6) * RW;
            T v[RW];
#pragma unroll
            for (int rr = 0; rr < RW; rr++) {
                const int r = rw0 + rr;
                T val = lane < bs0 ? g[r * TRI_NP + lane] : T(0);
                if (bs1 > 0 && lane < bs0)
                    val -= ((red[r * NT + lane] + red[r * NT + 64 + lane]) + red[r * NT + 128 + lane]) +
                           red[r * NT + 192 + lane];
                v[rr] = val;
            }
            wave_upper_solve<T, RW>(sd, LDD, bs0, lane, v);
            if (lane < bs0) {
#pragma unroll
                for (int rr = 0; rr < RW; rr++) g[(rw0 + rr) * TRI_NP + lane] = v[rr];
            }
        }
        __syncthreads();
        if (tid < np) {
#pragma unroll
            for (int r = 0; r < SOLVE_RC; r++)
                if (r < rc) x[(int64_t)(r0 + r) * a.n + c1 + tid] = g[r * TRI_NP + tid];
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------
// Fused per-front sweeps for fronts without stored inverses and nf <= FRONT_MAXNF: one launch
// per level and sweep instead of gather / triangle / split-K product / sum.
//   forward:  gather the children's update vectors (the sums of k_fwd_gather), permute the
//             pivot part by the block pivoting, then z -= [L11; L21](:, j) z_j for
//             j = 0 .. np-1; y <- pivot part, this front's update vector <- the rest.
//   backward: s = y_piv - U12 x_upd, then back substitution with U11.
// k_*_front_w: one wave per front (nf <= 64), the front vector in registers, exchanges by
// shuffles.  k_*_front_b: one 256-thread workgroup per front, the front vector in LDS,
// diagonal blocks of 64 staged in LDS and solved by one wave, products with 16 loads in
// flight per thread.
// ------------------------------------------------------------------------------------------
constexpr int FRONT_MAXNF = 1024;

template <typename T>
__device__ __forceinline__ void fwd_front_w(const SolveArgs &a, const int32_t *__restrict__ ecolptr,
                                            const int32_t *__restrict__ eupd, int slot, int mat, int zc, int nz)
{
    const SnDev d = a.sn[a.list[slot]];
    const int nf = d.nf, np = d.np, c1 = d.c1;
    const T *F = reinterpret_cast<const T *>(a.store) + mat * a.store_stride + d.foff;
    const int32_t *prow = a.prow + mat * a.n_stride;
    const int p = threadIdx.x & 63;
    const bool on = p < nf;
    const int64_t vb = (int64_t)mat * a.nrhs;
    const T *w0 = reinterpret_cast<const T *>(a.w0) + vb * a.n;
    T *y = reinterpret_cast<T *>(a.y) + vb * a.n;
    T *uv = reinterpret_cast<T *>(a.uv) + vb * a.nupd;
    const int nrhs = (int)a.nrhs;
    const int q0 = on ? ecolptr[d.fcol + p] : 0, q1 = on ? ecolptr[d.fcol + p + 1] : 0;
    const int src = p < np ? prow[c1 + p] - c1 : p;   // V-row feeding factor position p
    for (int r0 = zc * SOLVE_RC; r0 < nrhs; r0 += nz * SOLVE_RC) {
        const int rc = min(SOLVE_RC, nrhs - r0);
        T v[SOLVE_RC];
#pragma unroll
        for (int r = 0; r < SOLVE_RC; r++)
            v[r] = (r < rc && p < np) ? w0[(int64_t)(r0 + r) * a.n + c1 + p] : T(0);
        for (int q = q0; q < q1; q++) {
            const int32_t off = eupd[q];
#pragma unroll
            for (int r = 0; r < SOLVE_RC; r++)
                if (r < rc) v[r] += uv[(int64_t)(r0 + r) * a.nupd + off];
        }
#pragma unroll
        for (int r = 0; r < SOLVE_RC; r++) v[r] = __shfl(v[r], src);
        const T *row = F + p;
        // columns of [L11; L21] in groups of 8: the 8 loads are issued before the 8 steps
        for (int j0 = 0; j0 < np; j0 += 8) {
            T lv[8];
#pragma unroll
            for (int q = 0; q < 8; q++) lv[q] = (on && p > j0 + q && j0 + q < np) ? row[(int64_t)(j0 + q) * nf] : T(0);
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int j = j0 + q;
                if (j >= np) break;
#pragma unroll
                for (int r = 0; r < SOLVE_RC; r++) v[r] -= lv[q] * lane_value(v[r], j);
            }
        }
        if (on) {
#pragma unroll
            for (int r = 0; r < SOLVE_RC; r++) {
                if (r >= rc) continue;
                if (p < np) y[(int64_t)(r0 + r) * a.n + c1 + p] = v[r];
                else uv[(int64_t)(r0 + r) * a.nupd + d.upd_off + p - np] = v[r];
            }
        }
    }
}

template <typename T>
__device__ __forceinline__ void bwd_front_w(const SolveArgs &a, int slot, int mat, int zc, int nz)
{
    const SnDev d = a.sn[a.list[slot]];
    const int nf = d.nf, np = d.np, nu = d.nu, c1 = d.c1;
    const T *F = reinterpret_cast<const T *>(a.store) + mat * a.store_stride + d.foff;
    const int p = threadIdx.x & 63;
    const int64_t vb = (int64_t)mat * a.nrhs;
    const T *y = reinterpret_cast<const T *>(a.y) + vb * a.n;
    T *x = reinterpret_cast<T *>(a.x) + vb * a.n;
    const int nrhs = (int)a.nrhs;
    const int64_t src = p < np ? c1 + p : (p < nf ? a.upd[d.upd_off + p - np] : 0);
    for (int r0 = zc * SOLVE_RC; r0 < nrhs; r0 += nz * SOLVE_RC) {
        const int rc = min(SOLVE_RC, nrhs - r0);
        T v[SOLVE_RC];
#pragma unroll
        for (int r = 0; r < SOLVE_RC; r++)
            v[r] = (r < rc && p < nf) ? (p < np ? y : x)[(int64_t)(r0 + r) * a.n + src] : T(0);
        // s_i -= sum_t U12[i, t] x_upd[t] (the update values sit in lanes np + t)
        T acc[SOLVE_RC] = {};
        const T *urow = reinterpret_cast<const T *>(a.store) + mat * a.store_stride + d.u12off + p;   // U12, ld np
        for (int t0 = 0; t0 < nu; t0 += 8) {
            T uu[8];
#pragma unroll
            for (int q = 0; q < 8; q++) uu[q] = (p < np && t0 + q < nu) ? urow[(int64_t)(t0 + q) * np] : T(0);
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int t = t0 + q;
                if (t >= nu) break;
#pragma unroll
                for (int r = 0; r < SOLVE_RC; r++) acc[r] += uu[q] * lane_value(v[r], np + t);
            }
        }
#pragma unroll
        for (int r = 0; r < SOLVE_RC; r++) v[r] -= acc[r];
        const T *col = F + p;
        for (int j1 = np - 1; j1 >= 0; j1 -= 8) {
            T uc[8], dg[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int j = j1 - q;
                uc[q] = (j >= 0 && p < j) ? col[(int64_t)j * nf] : T(0);
                dg[q] = j >= 0 ? F[j + (int64_t)j * nf] : T(1);
            }
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int j = j1 - q;
                if (j < 0) break;
                const T rjj = T(1) / dg[q];   // one division per step
#pragma unroll
                for (int r = 0; r < SOLVE_RC; r++) {
                    const T t = lane_value(v[r], j) * rjj;
                    v[r] = p == j ? t : v[r] - uc[q] * t;
                }
            }
        }
        if (p < np) {
#pragma unroll
            for (int r = 0; r < SOLVE_RC; r++)
                if (r < rc) x[(int64_t)(r0 + r) * a.n + c1 + p] = v[r];
        }
    }
}

// The one-wave sweeps as grid-stride loops: each wave of a 256-thread workgroup takes work
// items (front, matrix, right-hand-side chunk) in turn, so a level of thousands of tiny fronts
// times a batch of matrices is a few thousand workgroups, not millions.
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_front_w(SolveArgs a, const int32_t *__restrict__ ecolptr,
                                                     const int32_t *__restrict__ eupd, int count, int nmat, int nz)
{
    const int64_t items = (int64_t)count * nmat * nz;
    for (int64_t it = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); it < items; it += (int64_t)gridDim.x * 4) {
        const int slot = (int)(it % count);
        const int64_t rest = it / count;
        fwd_front_w<T>(a, ecolptr, eupd, slot, (int)(rest % nmat), (int)(rest / nmat), nz);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_bwd_front_w(SolveArgs a, int count, int nmat, int nz)
{
    const int64_t items = (int64_t)count * nmat * nz;
    for (int64_t it = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); it < items; it += (int64_t)gridDim.x * 4) {
        const int slot = (int)(it % count);
        const int64_t rest = it / count;
        bwd_front_w<T>(a, slot, (int)(rest % nmat), (int)(rest / nmat), nz);
    }
}

template <typename T, int MAXNF>
__device__ __forceinline__ void fwd_front_b(const SolveArgs &a, const int32_t *__restrict__ ecolptr,
                                            const int32_t *__restrict__ eupd, int slot, int mat_, int zc, int nz)
{
    constexpr int NT = 256, PT = MAXNF / NT, LDD = SOLVE_NB + 1, DPT = SOLVE_NB * SOLVE_NB / NT;
    __shared__ T g[SOLVE_RC * MAXNF];            // g[r * MAXNF + p]
    __shared__ T sd[SOLVE_NB * LDD];             // diagonal block of L11
    const int mat = mat_;
    const SnDev d = a.sn[a.list[slot]];
    const int nf = d.nf, np = d.np, c1 = d.c1;
    const T *F = reinterpret_cast<const T *>(a.store) + mat * a.store_stride + d.foff;
    const int32_t *prow = a.prow + mat * a.n_stride;
    const int tid = threadIdx.x, lane = tid & 63;
    const int64_t vb = (int64_t)mat * a.nrhs;
    const T *w0 = reinterpret_cast<const T *>(a.w0) + vb * a.n;
    T *y = reinterpret_cast<T *>(a.y) + vb * a.n;
    T *uv = reinterpret_cast<T *>(a.uv) + vb * a.nupd;
    const int nrhs = (int)a.nrhs;
    // the extend-add ranges of this thread's positions p = tid + k NT
    int qb[PT], qe[PT];
    int nstep = 0;
#pragma unroll
    for (int k = 0; k < PT; k++) {
        const int p = tid + k * NT;
        qb[k] = p < nf ? ecolptr[d.fcol + p] : 0;
        qe[k] = p < nf ? ecolptr[d.fcol + p + 1] : 0;
        nstep = max(nstep, qe[k] - qb[k]);
    }
    for (int r0 = zc * SOLVE_RC; r0 < nrhs; r0 += nz * SOLVE_RC) {
        const int rc = min(SOLVE_RC, nrhs - r0);
        // the first diagonal block does not depend on the gather: its loads go out first
        T dpre[DPT];
        {
            const int bs = min(SOLVE_NB, np);
#pragma unroll
            for (int h = 0; h < DPT; h++) {
                const int e = tid + h * NT, i = e % SOLVE_NB, j = e / SOLVE_NB;
                dpre[h] = (i < bs && j < bs) ? F[i + (int64_t)j * nf] : T(0);
            }
        }
        // gather: own entries plus the children's update rows, child by child in list order,
        // every position of the thread at once
        T acc[PT][SOLVE_RC];
#pragma unroll
        for (int k = 0; k < PT; k++) {
            const int p = tid + k * NT;
#pragma unroll
            for (int r = 0; r < SOLVE_RC; r++)
                acc[k][r] = (r < rc && p < np) ? w0[(int64_t)(r0 + r) * a.n + c1 + p] : T(0);
        }
        for (int st = 0; st < nstep; st++) {
            int off[PT];
#pragma unroll
            for (int k = 0; k < PT; k++) off[k] = qb[k] + st < qe[k] ? eupd[qb[k] + st] : -1;
#pragma unroll
            for (int k = 0; k < PT; k++)
                if (off[k] >= 0) {
#pragma unroll
                    for (int r = 0; r < SOLVE_RC; r++)
                        if (r < rc) acc[k][r] += uv[(int64_t)(r0 + r) * a.nupd + off[k]];
                }
        }
#pragma unroll
        for (int k = 0; k < PT; k++) {
            const int p = tid + k * NT;
            if (p < nf) {
#pragma unroll
                for (int r = 0; r < SOLVE_RC; r++) g[r * MAXNF + p] = acc[k][r];
            }
        }
        __syncthreads();
        // pivot part into factor order (the block pivoting), through registers
#pragma unroll
        for (int k = 0; k < PT; k++) {
            const int i = tid + k * NT;
            if (i < np) {
                const int src = prow[c1 + i] - c1;
#pragma unroll
                for (int r = 0; r < SOLVE_RC; r++) acc[k][r] = g[r * MAXNF + src];
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PT; k++) {
            const int i = tid + k * NT;
            if (i < np) {
#pragma unroll
                for (int r = 0; r < SOLVE_RC; r++) g[r * MAXNF + i] = acc[k][r];
            }
        }
        for (int b0 = 0; b0 < np; b0 += SOLVE_NB) {
            const int bs = min(SOLVE_NB, np - b0);
            if (b0 > 0) {
#pragma unroll
                for (int h = 0; h < DPT; h++) {
                    const int e = tid + h * NT, i = e % SOLVE_NB, j = e / SOLVE_NB;
                    dpre[h] = (i < bs && j < bs) ? F[(b0 + i) + (int64_t)(b0 + j) * nf] : T(0);
                }
            }
#pragma unroll
            for (int h = 0; h < DPT; h++) {
                const int e = tid + h * NT;
                sd[(e % SOLVE_NB) + (e / SOLVE_NB) * LDD] = dpre[h];
            }
            __syncthreads();
            {
                // the right-hand sides split over the four waves, RW each
                const int rw0 = (tid >> 6) * RW;
                T v[RW];
#pragma unroll
                for (int rr = 0; rr < RW; rr++) {
                    const int r = rw0 + rr;
                    v[rr] = lane < bs ? g[r * MAXNF + b0 + lane] : T(0);
                }
                wave_lower_solve<T, RW>(sd, LDD, bs, lane, v);
                if (lane < bs) {
#pragma unroll
                    for (int rr = 0; rr < RW; rr++) {
                        const int r = rw0 + rr;
                        g[r * MAXNF + b0 + lane] = v[rr];
                    }
                }
            }
            __syncthreads();
            for (int p = b0 + bs + tid; p < nf; p += NT) {
                T acc1[SOLVE_RC] = {};
                dot_rows<T, 32>(F + p + (int64_t)b0 * nf, nf, bs, g + b0, MAXNF, acc1);
#pragma unroll
                for (int r = 0; r < SOLVE_RC; r++) g[r * MAXNF + p] -= acc1[r];
            }
            __syncthreads();
        }
        for (int p = tid; p < nf; p += NT) {
#pragma unroll
            for (int r = 0; r < SOLVE_RC; r++) {
                if (r >= rc) continue;
                if (p < np) y[(int64_t)(r0 + r) * a.n + c1 + p] = g[r * MAXNF + p];
                else uv[(int64_t)(r0 + r) * a.nupd + d.upd_off + p - np] = g[r * MAXNF + p];
            }
        }
        __syncthreads();
    }
}

template <typename T, int MAXNF>
__global__ __launch_bounds__(256) void k_fwd_front_b(SolveArgs a, const int32_t *__restrict__ ecolptr,
                                                     const int32_t *__restrict__ eupd)
{
    fwd_front_b<T, MAXNF>(a, ecolptr, eupd, blockIdx.x, blockIdx.y, blockIdx.z, gridDim.z);
}

template <typename T, int MAXNF>
__device__ __forceinline__ void bwd_front_b(const SolveArgs &a, int slot, int mat_, int zc, int nz)
{
    constexpr int NT = 256, LDD = SOLVE_NB + 1;
    constexpr int SDN = SOLVE_NB * LDD > NT * SOLVE_RC ? SOLVE_NB * LDD : NT * SOLVE_RC;
    __shared__ T g[SOLVE_RC * MAXNF];            // s at [0, np), x of the update rows at np + t
    __shared__ T sdred[SDN];                     // U12 product partial sums, then diagonal blocks of U11
    T *const sd = sdred;
    T *const red = sdred;
    const int mat = mat_;
    const SnDev d = a.sn[a.list[slot]];
    const int nf = d.nf, np = d.np, nu = d.nu, c1 = d.c1;
    const T *F = reinterpret_cast<const T *>(a.store) + mat * a.store_stride + d.foff;
    const T *U12 = reinterpret_cast<const T *>(a.store) + mat * a.store_stride + d.u12off;   // ld np
    const int32_t *upd = a.upd + d.upd_off;
    const int tid = threadIdx.x, lane = tid & 63;
    const int64_t vb = (int64_t)mat * a.nrhs;
    const T *y = reinterpret_cast<const T *>(a.y) + vb * a.n;
    T *x = reinterpret_cast<T *>(a.x) + vb * a.n;
    const int nrhs = (int)a.nrhs;
    // U12 product layout: thread (i, ts) with i = tid % npad over rows, ts = tid / npad over
    // columns t = ts, ts + ts_n, ...; the ts_n partial sums of a row are added in ts order
    int npad = 1;
    while (npad < np && npad < NT) npad <<= 1;
    const int ts_n = NT / npad, ti = tid % npad, ts = tid / npad;
    constexpr int PT = MAXNF / NT;
    int64_t srcs[PT];
#pragma unroll
    for (int k = 0; k < PT; k++) {
        const int p = tid + k * NT;
        srcs[k] = p < np ? c1 + p : (p < nf ? upd[p - np] : 0);
    }
    for (int r0 = zc * SOLVE_RC; r0 < nrhs; r0 += nz * SOLVE_RC) {
        const int rc = min(SOLVE_RC, nrhs - r0);
        // s = y of the pivot rows, x at the update columns: every position of the thread at once
        T gv[PT][SOLVE_RC];
#pragma unroll
        for (int k = 0; k < PT; k++) {
            const int p = tid + k * NT;
#pragma unroll
            for (int r = 0; r < SOLVE_RC; r++)
                gv[k][r] = (r < rc && p < nf) ? (p < np ? y : x)[(int64_t)(r0 + r) * a.n + srcs[k]] : T(0);
        }
#pragma unroll
        for (int k = 0; k < PT; k++) {
            const int p = tid + k * NT;
            if (p < nf) {
#pragma unroll
                for (int r = 0; r < SOLVE_RC; r++) g[r * MAXNF + p] = gv[k][r];
            }
        }
        __syncthreads();
        for (int i0 = 0; i0 < np; i0 += npad) {
            const int i = i0 + ti;
            T acc[SOLVE_RC] = {};
            if (i < np) {
                const T *row = U12 + i + (int64_t)ts * np;
                const int cnt = nu > ts ? (nu - ts + ts_n - 1) / ts_n : 0;   // columns ts + k * ts_n
                int k = 0;
                for (; k + 16 <= cnt; k += 16) {
                    T u[16];
#pragma unroll
                    for (int q = 0; q < 16; q++) u[q] = row[(int64_t)(k + q) * ts_n * np];
#pragma unroll
                    for (int q = 0; q < 16; q++)
#pragma unroll
                        for (int r = 0; r < SOLVE_RC; r++) acc[r] += u[q] * g[r * MAXNF + np + ts + (k + q) * ts_n];
                }
                for (; k + 4 <= cnt; k += 4) {
                    T u[4];
#pragma unroll
                    for (int q = 0; q < 4; q++) u[q] = row[(int64_t)(k + q) * ts_n * np];
#pragma unroll
                    for (int q = 0; q < 4; q++)
#pragma unroll
                        for (int r = 0; r < SOLVE_RC; r++) acc[r] += u[q] * g[r * MAXNF + np + ts + (k + q) * ts_n];
                }
                for (; k < cnt; k++) {
                    const T u = row[(int64_t)k * ts_n * np];
#pragma unroll
                    for (int r = 0; r < SOLVE_RC; r++) acc[r] += u * g[r * MAXNF + np + ts + k * ts_n];
                }
            }
#pragma unroll
            for (int r = 0; r < SOLVE_RC; r++) red[r * NT + tid] = acc[r];
            __syncthreads();
            // fixed binary tree over the ts_n partial sums of each row (log2(ts_n) steps)
            for (int half = ts_n >> 1; half > 0; half >>= 1) {
                if (ts < half) {
#pragma unroll
                    for (int r = 0; r < SOLVE_RC; r++) red[r * NT + tid] += red[r * NT + tid + half * npad];
                }
                __syncthreads();
            }
            if (tid < npad && i0 + tid < np) {
#pragma unroll
                for (int r = 0; r < SOLVE_RC; r++) g[r * MAXNF + i0 + tid] -= red[r * NT + tid];
            }
            __syncthreads();
        }
        for (int b0 = ((np - 1) / SOLVE_NB) * SOLVE_NB; b0 >= 0; b0 -= SOLVE_NB) {
            const int bs = min(SOLVE_NB, np - b0);
            for (int e = tid; e < SOLVE_NB * SOLVE_NB; e += NT) {
                const int i = e % SOLVE_NB, j = e / SOLVE_NB;
                sd[i + j * LDD] = (i < bs && j < bs) ? F[(b0 + i) + (int64_t)(b0 + j) * nf] : T(0);
            }
            __syncthreads();
            {
                // the right-hand sides split over the four waves, RW each
                const int rw0 = (tid >> 6) * RW;
                T v[RW];
#pragma unroll
                for (int rr = 0; rr < RW; rr++) {
                    const int r = rw0 + rr;
                    v[rr] = lane < bs ? g[r * MAXNF + b0 + lane] : T(0);
                }
                wave_upper_solve<T, RW>(sd, LDD, bs, lane, v);
                if (lane < bs) {
#pragma unroll
                    for (int rr = 0; rr < RW; rr++) {
                        const int r = rw0 + rr;
                        g[r * MAXNF + b0 + lane] = v[rr];
                    }
                }
            }
            __syncthreads();
            for (int i = tid; i < b0; i += NT) {
                T acc[SOLVE_RC] = {};
                dot_rows<T, 16>(F + i + (int64_t)b0 * nf, nf, bs, g + b0, MAXNF, acc);
#pragma unroll
                for (int r = 0; r < SOLVE_RC; r++) g[r * MAXNF + i] -= acc[r];
            }
            __syncthreads();
        }
        for (int i = tid; i < np; i += NT) {
#pragma unroll
            for (int r = 0; r < SOLVE_RC; r++)
                if (r < rc) x[(int64_t)(r0 + r) * a.n + c1 + i] = g[r * MAXNF + i];
        }
        __syncthreads();
    }
}

template <typename T, int MAXNF>
__global__ __launch_bounds__(256) void k_bwd_front_b(SolveArgs a)
{
    bwd_front_b<T, MAXNF>(a, blockIdx.x, blockIdx.y, blockIdx.z, gridDim.z);
}

__global__ __launch_bounds__(256) void k_copy16(int4 *__restrict__ dst, const int4 *__restrict__ src, int64_t n16)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

hipError_t launch_copy(void *dst, const void *src, int64_t bytes, hipStream_t st)
{
    const int64_t n16 = bytes / 16;
    const unsigned grid = (unsigned)std::min<int64_t>((n16 + 255) / 256, 8192);
    hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, st, reinterpret_cast<int4 *>(dst),
                       reinterpret_cast<const int4 *>(src), n16);
    return hipGetLastError();
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

// R = B - A X with A in CSR: one thread per (row, chunk of eight right-hand sides) of one
// matrix, the row's entries read once per chunk; each R(i, c) sums its row in entry order
// from 0.  Grid x = row blocks times chunks, y = matrix.
constexpr int RES_RC = 8;
__global__ __launch_bounds__(256) void k_residual(const int32_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                                  const int32_t *__restrict__ vmap, const double *__restrict__ ax,
                                                  int64_t ax_stride, const double *__restrict__ X, int64_t ldx,
                                                  const double *__restrict__ B, int64_t ldb, double *__restrict__ R,
                                                  int64_t n, int64_t nrhs, int64_t mat_stride_x, int64_t mat_stride_b)
{
    const int mat = blockIdx.y;
    const int64_t nrb = (n + 255) / 256;
    const int64_t i = (int64_t)(blockIdx.x % nrb) * 256 + threadIdx.x;
    const int64_t c0 = (int64_t)(blockIdx.x / nrb) * RES_RC;
    if (i >= n || c0 >= nrhs) return;
    const double *axm = ax + mat * ax_stride;
    const double *Xm = X + mat * mat_stride_x;
    const int p0 = rp[i], p1 = rp[i + 1];
    const int rc = (int)min((int64_t)RES_RC, nrhs - c0);
    double s[RES_RC];
#pragma unroll
    for (int c = 0; c < RES_RC; c++) s[c] = 0.0;
    if (rc == RES_RC) {   // full chunk: the eight loads of an entry go out together
        for (int p = p0; p < p1; p++) {
            const double a = axm[vmap[p]];
            const double *xj = Xm + ci[p] + c0 * ldx;
            double xv[RES_RC];
#pragma unroll
            for (int c = 0; c < RES_RC; c++) xv[c] = xj[c * ldx];
#pragma unroll
            for (int c = 0; c < RES_RC; c++) s[c] += a * xv[c];
        }
    } else {
        for (int p = p0; p < p1; p++) {
            const double a = axm[vmap[p]];
            const double *xj = Xm + ci[p] + c0 * ldx;
#pragma unroll
            for (int c = 0; c < RES_RC; c++)
                if (c < rc) s[c] += a * xj[c * ldx];
        }
    }
#pragma unroll
    for (int c = 0; c < RES_RC; c++)
        if (c < rc) R[mat * mat_stride_b + i + (c0 + c) * n] = B[mat * mat_stride_b + i + (c0 + c) * ldb] - s[c];
}

// Norms for the backward error, two passes.  Pass 1: grid (nb, nrhs, nmat), nb <= NORM_BLOCKS
// blocks per column (about 2048 entries each); block b of column c reduces its slice to
// part[((mat * nrhs + c) * nb + b) * 5 + q] with q = 0 max|R|, 1 max|X|, 2 max|B|, 3 sum R^2,
// 4 sum B^2.  Pass 2 folds the blocks in order.
constexpr int NORM_BLOCKS = 64;

__global__ __launch_bounds__(256) void k_norms_part(const double *__restrict__ R, const double *__restrict__ X,
                                                    const double *__restrict__ B, int64_t n, int64_t nrhs,
                                                    double *__restrict__ part)
{
    const int b = blockIdx.x, col = blockIdx.y, mat = blockIdx.z, nb = gridDim.x;
    const int64_t off = ((int64_t)mat * nrhs + col) * n;
    double mr = 0, mx = 0, mb = 0, sr = 0, sb = 0;
    for (int64_t i = (int64_t)b * blockDim.x + threadIdx.x; i < n; i += (int64_t)nb * blockDim.x) {
        const double rv = R[off + i], xv = X[off + i], bv = B[off + i];
        mr = fmax(mr, fabs(rv)); mx = fmax(mx, fabs(xv)); mb = fmax(mb, fabs(bv));
        sr += rv * rv; sb += bv * bv;
    }
    __shared__ double sh[5][256];
    sh[0][threadIdx.x] = mr; sh[1][threadIdx.x] = mx; sh[2][threadIdx.x] = mb;
    sh[3][threadIdx.x] = sr; sh[4][threadIdx.x] = sb;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            for (int q = 0; q < 3; q++) sh[q][threadIdx.x] = fmax(sh[q][threadIdx.x], sh[q][threadIdx.x + st]);
            for (int q = 3; q < 5; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x < 5) part[(((int64_t)mat * nrhs + col) * nb + b) * 5 + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void k_norms_fold(const double *__restrict__ part, int64_t ncols, int nb, double *__restrict__ out)
{
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= ncols) return;
    double v[5] = {0, 0, 0, 0, 0};
    for (int b = 0; b < nb; b++) {
        const double *p = part + (c * nb + b) * 5;
        for (int q = 0; q < 3; q++) v[q] = fmax(v[q], p[q]);
        for (int q = 3; q < 5; q++) v[q] += p[q];
    }
    for (int q = 0; q < 5; q++) out[c * 5 + q] = v[q];
}

// ||A||_inf per matrix (max row sum of |a_ij|): per-block maxima, then an in-order fold.
__global__ __launch_bounds__(256) void k_anorm_part(const int32_t *__restrict__ rp, const int32_t *__restrict__ vmap,
                                                    const double *__restrict__ ax, int64_t ax_stride, int64_t n,
                                                    double *__restrict__ part)
{
    const double *axm = ax + blockIdx.y * ax_stride;
    double m = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double s = 0;
        for (int p = rp[i]; p < rp[i + 1]; p++) s += fabs(axm[vmap[p]]);
        m = fmax(m, s);
    }
    __shared__ double sh[256];
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = sh[0];
}

__global__ void k_anorm_fold(const double *__restrict__ part, int nblk, int nmat, double *__restrict__ out)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmat) return;
    double v = 0;
    for (int b = 0; b < nblk; b++) v = fmax(v, part[m * nblk + b]);
    out[m] = v;
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
__global__ __launch_bounds__(256) void k_export(const SnDev *__restrict__ sn, const T *__restrict__ store,
                                                const int32_t *__restrict__ upd, const int32_t *__restrict__ rowfac,
                                                const int64_t *__restrict__ Lp, const int64_t *__restrict__ Up,
                                                const int64_t *__restrict__ uoff, int64_t *__restrict__ Li,
                                                double *__restrict__ Lx, int64_t *__restrict__ Ui,
                                                double *__restrict__ Ux)
{
    const SnDev d = sn[blockIdx.x];
    const int nf = d.nf, np = d.np, nu = d.nu, c1 = d.c1;
    const T *F = store + d.foff, *U12 = store + d.u12off;
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
            Ux[up + b] = (double)U12[b + (int64_t)t * np];
        }
    }
}

// ------------------------------------------------------------------------------------------
// Column-parallel sweeps, for solves with many right-hand-side columns (a batch of matrices
// times their right-hand sides).  Column g = mat * nrhs + r of the batch is one lane; a wave
// holds 64 consecutive columns and a work item is (front, wave of columns).  Each lane runs
// the front's substitution for its own column in the order of the one-column sweeps:
//   forward:  z = own entries + the children's update rows (extend-add list order), pivot part
//             permuted by the block pivoting, y_j = z_j and z_i -= L(i, j) y_j for j ascending;
//   backward: s = y_piv - U12 x_upd, then x_j = s_j / U(j, j) and s_i -= U(i, j) x_j for j
//             descending.
// The vectors are interleaved by column (entry i of column g at i * C + g, C = nmat * nrhs),
// so each vector access of a wave is 64 consecutive words.  When nrhs is a multiple of 64 a
// wave holds one matrix and every factor entry is one scalar load for the 64 lanes.  Rows
// pass through registers in chunks of CH <= 64; factor entries of a chunk beyond the front are
// read from the store's slack (64 elements past each matrix's factors) and never written back.
// ------------------------------------------------------------------------------------------
// z[r] -= col_j[r] * v_j for the columns j in [j0, j1) of a chunk, four columns per step:
// the four vector values are loaded before any of their products
// z[r] -= col_j[r] * v_j for the columns j in [j0, j1) of a chunk, eight columns per step,
// their vector values and factor columns loaded before any of their products.  UNI (the
// wave's lanes share one matrix): lane l loads row l of each column segment with one vector
// load and row r reaches every lane by v_readlane, so eight columns are in flight without
// tying up scalar registers; otherwise each lane reads its own matrix's entries.
template <typename T, int CH, bool UNI, typename VF, typename CF>
__device__ __forceinline__ void cs_cols(T (&z)[CH], int j0, int j1, VF vat, CF colat)
{
    const int lane = threadIdx.x & 63;
    int j = j0;
    for (; j + 8 <= j1; j += 8) {
        T v[8];
        if constexpr (UNI) {
            T f[8];
#pragma unroll
            for (int q = 0; q < 8; q++) f[q] = colat(j + q)[lane];
#pragma unroll
            for (int q = 0; q < 8; q++) v[q] = vat(j + q);
#pragma unroll
            for (int q = 0; q < 8; q++)
#pragma unroll
                for (int r = 0; r < CH; r++) z[r] -= lane_value(f[q], r) * v[q];
        } else {
#pragma unroll
            for (int q = 0; q < 8; q++) v[q] = vat(j + q);
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const T *col = colat(j + q);
#pragma unroll
                for (int r = 0; r < CH; r++) z[r] -= col[r] * v[q];
            }
        }
    }
    for (; j < j1; j++) {
        const T vj = vat(j);
        if constexpr (UNI) {
            const T f = colat(j)[lane];
#pragma unroll
            for (int r = 0; r < CH; r++) z[r] -= lane_value(f, r) * vj;
        } else {
            const T *col = colat(j);
#pragma unroll
            for (int r = 0; r < CH; r++) z[r] -= col[r] * vj;
        }
    }
}

// the chunk's own unit lower triangle: y_j = z_j, z_r -= L(r, j) y_j for r > j
template <typename T, int CH, bool UNI>
__device__ __forceinline__ void cs_lower(T (&z)[CH], const T *D, int ld, int nb)
{
    const int lane = threadIdx.x & 63;
    if constexpr (UNI) {
        T f[CH];
#pragma unroll
        for (int jj = 0; jj < CH; jj++) f[jj] = jj < nb ? D[(int64_t)jj * ld + lane] : T(0);
#pragma unroll
        for (int jj = 0; jj < CH; jj++) {
            if (jj >= nb) continue;
            const T yj = z[jj];
#pragma unroll
            for (int r = jj + 1; r < CH; r++) z[r] -= lane_value(f[jj], r) * yj;
        }
    } else {
#pragma unroll
        for (int jj = 0; jj < CH; jj++) {
            if (jj >= nb) continue;
            const T yj = z[jj];
            const T *col = D + (int64_t)jj * ld;
#pragma unroll
            for (int r = jj + 1; r < CH; r++) z[r] -= col[r] * yj;
        }
    }
}

// the chunk's upper triangle, last column first: x_j = z_j / U(j, j), z_r -= U(r, j) x_j for r < j
template <typename T, int CH, bool UNI>
__device__ __forceinline__ void cs_upper(T (&z)[CH], const T *D, int ld, int nb)
{
    const int lane = threadIdx.x & 63;
    if constexpr (UNI) {
        T f[CH];
#pragma unroll
        for (int jj = 0; jj < CH; jj++) f[jj] = jj < nb ? D[(int64_t)jj * ld + lane] : T(1);
#pragma unroll
        for (int jj = CH - 1; jj >= 0; jj--) {
            if (jj >= nb) continue;
            const T xj = z[jj] / lane_value(f[jj], jj);
            z[jj] = xj;
#pragma unroll
            for (int r = 0; r < jj; r++) z[r] -= lane_value(f[jj], r) * xj;
        }
    } else {
#pragma unroll
        for (int jj = CH - 1; jj >= 0; jj--) {
            if (jj >= nb) continue;
            const T *col = D + (int64_t)jj * ld;
            const T xj = z[jj] / col[jj];
            z[jj] = xj;
#pragma unroll
            for (int r = 0; r < jj; r++) z[r] -= col[r] * xj;
        }
    }
}

// z[r] = base(r) + the children's update rows of front positions p0 + r (r < nb): the
// extend-add entries of those positions are one contiguous range of the lists, walked eight
// at a time with their loads in flight and summed per position in list order through this
// wave's LDS rows acc[r * 64 + lane].
template <typename T, int CH, typename BF>
__device__ __forceinline__ void cs_gather(T (&z)[CH], const SolveArgs &a, const SnDev &d, int p0, int nb, int64_t C,
                                          int64_t gc, T *acc, BF base)
{
    const int lane = threadIdx.x & 63;
    const T *uvr = reinterpret_cast<const T *>(a.uv);
#pragma unroll
    for (int r = 0; r < CH; r++) acc[r * 64 + lane] = r < nb ? base(r) : T(0);
    const int qa = a.ecolptr[d.fcol + p0], qb = a.ecolptr[d.fcol + p0 + nb];
    const int pbase = d.fcol + p0;
    for (int q = qa; q < qb; q += 8) {
        int off[8], pos[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int qq = min(q + k, qb - 1);
            off[k] = a.eupd[qq];
            pos[k] = a.egpos[qq] - pbase;
        }
        T v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = uvr[(int64_t)off[k] * C + gc];
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (q + k < qb) acc[pos[k] * 64 + lane] += v[k];
    }
#pragma unroll
    for (int r = 0; r < CH; r++) z[r] = acc[r * 64 + lane];
}

template <typename T, bool UNI, int CH>
__device__ __forceinline__ void cs_fwd_front(const SolveArgs &a, int slot, int64_t gw, int64_t C, int nrhs, T *acc)
{
    const int lane = threadIdx.x & 63;
    const int64_t g = gw + lane;
    const bool on = g < C;
    const int64_t gc = on ? g : C - 1;                 // loads of idle lanes stay in bounds
    const int mat = UNI ? (int)(gw / nrhs) : (int)(gc / nrhs);
    const SnDev d = a.sn[a.list[slot]];
    const int nf = d.nf, np = d.np, c1 = d.c1;
    const T *F = reinterpret_cast<const T *>(a.store) + (int64_t)mat * a.store_stride + d.foff;
    const int32_t *prow = a.prow + (int64_t)mat * a.n_stride;
    const T *w0 = reinterpret_cast<const T *>(a.w0);
    T *y = reinterpret_cast<T *>(a.y);
    T *uv = reinterpret_cast<T *>(a.uv);
    T *zs = reinterpret_cast<T *>(a.zbuf);            // the gathered pivot part, V-row order
    const auto yat = [&](int j) { return y[(int64_t)(c1 + j) * C + gc]; };
    // gather the pivot part (own entries plus the children's update rows), chunk by chunk
    for (int p0 = 0; p0 < np; p0 += CH) {
        const int nb = min(CH, np - p0);
        T z[CH];
        cs_gather<T, CH>(z, a, d, p0, nb, C, gc, acc, [&](int r) { return w0[(int64_t)(c1 + p0 + r) * C + gc]; });
        if (on) {
#pragma unroll
            for (int r = 0; r < CH; r++)
                if (r < nb) zs[(int64_t)(c1 + p0 + r) * C + g] = z[r];
        }
    }
    // pivot part, chunk by chunk: earlier chunks' y, then the chunk's own unit lower triangle
    for (int b0 = 0; b0 < np; b0 += CH) {
        const int nb = min(CH, np - b0);
        T z[CH];
#pragma unroll
        for (int r = 0; r < CH; r++)   // factor position b0 + r takes V-row prow[c1 + b0 + r]
            z[r] = r < nb ? zs[(int64_t)(prow[c1 + b0 + r]) * C + gc] : T(0);
        cs_cols<T, CH, UNI>(z, 0, b0, yat, [&](int j) { return F + (int64_t)j * nf + b0; });
        cs_lower<T, CH, UNI>(z, F + (int64_t)b0 * nf + b0, nf, nb);
        if (on) {
#pragma unroll
            for (int r = 0; r < CH; r++)
                if (r < nb) y[(int64_t)(c1 + b0 + r) * C + g] = z[r];
        }
    }
    // update rows: the children's contributions minus L21 y
    for (int u0 = np; u0 < nf; u0 += CH) {
        const int nb = min(CH, nf - u0);
        T z[CH];
        cs_gather<T, CH>(z, a, d, u0, nb, C, gc, acc, [&](int) { return T(0); });
        cs_cols<T, CH, UNI>(z, 0, np, yat, [&](int j) { return F + (int64_t)j * nf + u0; });
        if (on) {
#pragma unroll
            for (int r = 0; r < CH; r++)
                if (r < nb) uv[(int64_t)(d.upd_off + u0 - np + r) * C + g] = z[r];
        }
    }
}

template <typename T, bool UNI, int CH>
__device__ __forceinline__ void cs_bwd_front(const SolveArgs &a, int slot, int64_t gw, int64_t C, int nrhs)
{
    const int lane = threadIdx.x & 63;
    const int64_t g = gw + lane;
    const bool on = g < C;
    const int64_t gc = on ? g : C - 1;
    const int mat = UNI ? (int)(gw / nrhs) : (int)(gc / nrhs);
    const SnDev d = a.sn[a.list[slot]];
    const int nf = d.nf, np = d.np, nu = d.nu, c1 = d.c1;
    const T *F = reinterpret_cast<const T *>(a.store) + (int64_t)mat * a.store_stride + d.foff;
    const T *U12 = reinterpret_cast<const T *>(a.store) + (int64_t)mat * a.store_stride + d.u12off;   // ld np
    const int32_t *upd = a.upd + d.upd_off;
    const T *y = reinterpret_cast<const T *>(a.y);
    T *x = reinterpret_cast<T *>(a.x);
    for (int b0 = ((np - 1) / CH) * CH; b0 >= 0; b0 -= CH) {
        const int nb = min(CH, np - b0);
        T z[CH];
#pragma unroll
        for (int r = 0; r < CH; r++) z[r] = r < nb ? y[(int64_t)(c1 + b0 + r) * C + gc] : T(0);
        cs_cols<T, CH, UNI>(z, 0, nu, [&](int t) { return x[(int64_t)upd[t] * C + gc]; },
                       [&](int t) { return U12 + (int64_t)t * np + b0; });
        cs_cols<T, CH, UNI>(z, b0 + CH, np, [&](int j) { return x[(int64_t)(c1 + j) * C + gc]; },
                       [&](int j) { return F + (int64_t)j * nf + b0; });
        cs_upper<T, CH, UNI>(z, F + (int64_t)b0 * nf + b0, nf, nb);
        if (on) {
#pragma unroll
            for (int r = 0; r < CH; r++)
                if (r < nb) x[(int64_t)(c1 + b0 + r) * C + g] = z[r];
        }
    }
}

// One level of a column-parallel sweep: items (front, wave of columns) in a grid-stride loop,
// four waves per workgroup; consecutive waves take the same front.  CH: rows per register
// chunk, chosen per level from its front sizes (small chunks keep more waves resident).
// Waves per SIMD the register budget must allow: the sweeps of small fronts are chains of
// dependent loads, hidden only by many resident waves.
template <typename T, int CH>
constexpr int cs_min_waves()
{
    return sizeof(T) == 4 ? (CH <= 16 ? 8 : (CH <= 32 ? 5 : 3)) : (CH <= 16 ? 5 : 3);
}

template <typename T, bool UNI, bool FWD, int CH>
__global__ __launch_bounds__(256, (cs_min_waves<T, CH>())) void k_cs_sweep(SolveArgs a, int count, int64_t C, int nrhs)
{
    __shared__ T acc[FWD ? 4 * CH * 64 : 1];   // per wave: the gather rows of one chunk
    const int64_t ncg = (C + 63) / 64;
    const int64_t items = (int64_t)count * ncg;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int64_t it = (int64_t)blockIdx.x * 4 + wave; it < items; it += (int64_t)gridDim.x * 4) {
        const int slot = (int)(it / ncg);
        const int64_t gw = (it - (int64_t)slot * ncg) * 64;
        if (FWD)
            cs_fwd_front<T, UNI, CH>(a, slot, gw, C, nrhs, acc + wave * CH * 64);
        else
            cs_bwd_front<T, UNI, CH>(a, slot, gw, C, nrhs);
    }
}

// Interleaved right-hand sides: w0[i * C + g] = Rs[qf[i]] * B[qf[i], r] of matrix mat (column g
// = mat * nrhs + r), through a 64 x 64 LDS tile: B is read down its columns (consecutive
// original rows q), w0 written along its rows (consecutive columns g) at i = qinv[q].
template <typename T>
__global__ __launch_bounds__(256) void k_cs_prep(const double *__restrict__ B, int64_t ldb, int64_t b_mat_stride,
                                                 const double *__restrict__ rs, int64_t rs_stride,
                                                 const int32_t *__restrict__ qinv, T *__restrict__ w0, int64_t n,
                                                 int nrhs, int64_t C)
{
    __shared__ double tile[64][65];
    const int64_t g0 = (int64_t)blockIdx.x * 64;
    for (int64_t q0 = (int64_t)blockIdx.y * 64; q0 < n; q0 += (int64_t)gridDim.y * 64) {
        for (int e = threadIdx.x; e < 64 * 64; e += 256) {
            const int ql = e & 63, gl = e >> 6;
            const int64_t g = g0 + gl, q = q0 + ql;
            double v = 0.0;
            if (g < C && q < n) {
                const int mat = (int)g / nrhs, r = (int)g - mat * nrhs;   // C < 2^31 (launcher)
                v = B[mat * b_mat_stride + q + r * ldb];
                if (rs) v *= rs[mat * rs_stride + q];
            }
            tile[gl][ql] = v;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < 64 * 64; e += 256) {
            const int gl = e & 63, ql = e >> 6;
            const int64_t g = g0 + gl, q = q0 + ql;
            if (g < C && q < n) w0[(int64_t)qinv[q] * C + g] = (T)tile[gl][ql];
        }
        __syncthreads();
    }
}

// X[qf[k], r] of matrix mat = x[k * C + g] (accumulate: +=, the FP64 refinement update), through
// the same tiles read along x's rows and written down X's columns.
template <typename T>
__global__ __launch_bounds__(256) void k_cs_finish(const T *__restrict__ x, const int32_t *__restrict__ qinv,
                                                   double *__restrict__ X, int64_t ldx, int64_t x_mat_stride,
                                                   int64_t n, int nrhs, int64_t C, int accumulate)
{
    __shared__ double tile[64][65];
    const int64_t g0 = (int64_t)blockIdx.x * 64;
    for (int64_t q0 = (int64_t)blockIdx.y * 64; q0 < n; q0 += (int64_t)gridDim.y * 64) {
        for (int e = threadIdx.x; e < 64 * 64; e += 256) {
            const int gl = e & 63, ql = e >> 6;
            const int64_t g = g0 + gl, q = q0 + ql;
            tile[gl][ql] = (g < C && q < n) ? (double)x[(int64_t)qinv[q] * C + g] : 0.0;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < 64 * 64; e += 256) {
            const int ql = e & 63, gl = e >> 6;
            const int64_t g = g0 + gl, q = q0 + ql;
            if (g < C && q < n) {
                const int mat = (int)g / nrhs, r = (int)g - mat * nrhs;   // C < 2^31 (launcher)
                double *dst = X + mat * x_mat_stride + q + r * ldx;
                *dst = accumulate ? *dst + tile[gl][ql] : tile[gl][ql];
            }
        }
        __syncthreads();
    }
}

// ---- host launchers ----
// A level's one-wave fronts and workgroup fronts in one launch: the first nb * nmat * nz
// workgroups take the larger fronts (front, matrix, right-hand-side chunk), the rest run the
// one-wave items of wlist in a grid-stride loop.
template <typename T, int MAXNF>
__global__ __launch_bounds__(256) void k_fwd_front_wb(SolveArgs a, const int32_t *__restrict__ ecolptr,
                                                      const int32_t *__restrict__ eupd, const int32_t *__restrict__ wlist,
                                                      int nb, int wcount, int nmat, int nz)
{
    const int nbw = nb * nmat * nz;
    if ((int)blockIdx.x < nbw) {
        const int rest = (int)blockIdx.x / nb;
        fwd_front_b<T, MAXNF>(a, ecolptr, eupd, (int)blockIdx.x % nb, rest % nmat, rest / nmat, nz);
        return;
    }
    SolveArgs aw = a;
    aw.list = wlist;
    const int64_t items = (int64_t)wcount * nmat * nz, nwg = (int64_t)gridDim.x - nbw;
    for (int64_t it = ((int64_t)blockIdx.x - nbw) * 4 + (threadIdx.x >> 6); it < items; it += nwg * 4) {
        const int64_t rest = it / wcount;
        fwd_front_w<T>(aw, ecolptr, eupd, (int)(it % wcount), (int)(rest % nmat), (int)(rest / nmat), nz);
    }
}

template <typename T, int MAXNF>
__global__ __launch_bounds__(256) void k_bwd_front_wb(SolveArgs a, const int32_t *__restrict__ wlist, int nb, int wcount,
                                                      int nmat, int nz)
{
    const int nbw = nb * nmat * nz;
    if ((int)blockIdx.x < nbw) {
        const int rest = (int)blockIdx.x / nb;
        bwd_front_b<T, MAXNF>(a, (int)blockIdx.x % nb, rest % nmat, rest / nmat, nz);
        return;
    }
    SolveArgs aw = a;
    aw.list = wlist;
    const int64_t items = (int64_t)wcount * nmat * nz, nwg = (int64_t)gridDim.x - nbw;
    for (int64_t it = ((int64_t)blockIdx.x - nbw) * 4 + (threadIdx.x >> 6); it < items; it += nwg * 4) {
        const int64_t rest = it / wcount;
        bwd_front_w<T>(aw, (int)(it % wcount), (int)(rest % nmat), (int)(rest / nmat), nz);
    }
}

template <typename T>
hipError_t launch_solve_merged(int fwd, int small, const SolveArgs &a, const int32_t *lists, int boff, int bcount,
                               int woff, int wcount, int nmat, hipStream_t st)
{
    SolveArgs b = a;
    b.list = lists + boff;
    const int nz = (int)((a.nrhs + SOLVE_RC - 1) / SOLVE_RC);
    const int64_t witems = (int64_t)wcount * nmat * nz;
    const int64_t grid = (int64_t)bcount * nmat * nz + std::min<int64_t>((witems + 3) / 4, 8192);
    if (grid > INT32_MAX) return hipErrorInvalidValue;
    const dim3 g((unsigned)grid), blk(256);
    const int32_t *wl = lists + woff;
    if (fwd) {
        if (small) hipLaunchKernelGGL((k_fwd_front_wb<T, 256>), g, blk, 0, st, b, a.ecolptr, a.eupd, wl, bcount, wcount, nmat, nz);
        else hipLaunchKernelGGL((k_fwd_front_wb<T, FRONT_MAXNF>), g, blk, 0, st, b, a.ecolptr, a.eupd, wl, bcount, wcount, nmat, nz);
    } else {
        if (small) hipLaunchKernelGGL((k_bwd_front_wb<T, 256>), g, blk, 0, st, b, wl, bcount, wcount, nmat, nz);
        else hipLaunchKernelGGL((k_bwd_front_wb<T, FRONT_MAXNF>), g, blk, 0, st, b, wl, bcount, wcount, nmat, nz);
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_solve_level(int which, const SolveArgs &a, const int32_t *lists, int list_off, int count, int nmat,
                              hipStream_t st)
{
    SolveArgs b = a;
    b.list = lists + list_off;
    // right-hand-side chunks of SOLVE_RC in grid z
    const int nz = (int)((a.nrhs + SOLVE_RC - 1) / SOLVE_RC);
    const bool sum = which == 2 || which == 4 || which == 7 || which == 9;   // RECT_SUM_TPR workgroups per item
    const bool prod = which == 1 || which == 3 || which == 6 || which == 8 || which == 17;   // two workgroups per item
    dim3 g(sum ? RECT_SUM_TPR * count : prod ? 2 * count : count, nmat, nz), blk(256);
    switch (which) {
    case 0: hipLaunchKernelGGL(k_fwd_tri<T>, g, blk, 0, st, b); break;
    case 1: hipLaunchKernelGGL(k_rect_part<T>, g, blk, 0, st, b, 0); break;
    case 2: hipLaunchKernelGGL(k_rect_sum<T>, g, blk, 0, st, b, 0); break;
    case 3: hipLaunchKernelGGL(k_rect_part<T>, g, blk, 0, st, b, 1); break;
    case 4: hipLaunchKernelGGL(k_rect_sum<T>, g, blk, 0, st, b, 1); break;
    case 5: hipLaunchKernelGGL(k_bwd_tri<T>, g, blk, 0, st, b); break;
    case 6: hipLaunchKernelGGL(k_rect_part<T>, g, blk, 0, st, b, 2); break;
    case 7: hipLaunchKernelGGL(k_rect_sum<T>, g, blk, 0, st, b, 2); break;
    case 8: hipLaunchKernelGGL(k_rect_part<T>, g, blk, 0, st, b, 3); break;
    case 9: hipLaunchKernelGGL(k_rect_sum<T>, g, blk, 0, st, b, 3); break;
    case 10: hipLaunchKernelGGL(k_fwd_gather<T>, g, blk, 0, st, b, a.ecolptr, a.eupd); break;
    case 11:
    case 13: {
        const int64_t items = (int64_t)count * nmat * nz;
        const unsigned wg = (unsigned)std::min<int64_t>((items + 3) / 4, 8192);
        if (which == 11)
            hipLaunchKernelGGL(k_fwd_front_w<T>, dim3(wg), blk, 0, st, b, a.ecolptr, a.eupd, count, nmat, nz);
        else
            hipLaunchKernelGGL(k_bwd_front_w<T>, dim3(wg), blk, 0, st, b, count, nmat, nz);
        break;
    }
    case 12: hipLaunchKernelGGL((k_fwd_front_b<T, FRONT_MAXNF>), g, blk, 0, st, b, a.ecolptr, a.eupd); break;
    case 14: hipLaunchKernelGGL((k_bwd_front_b<T, FRONT_MAXNF>), g, blk, 0, st, b); break;
    case 15: hipLaunchKernelGGL((k_fwd_front_b<T, 256>), g, blk, 0, st, b, a.ecolptr, a.eupd); break;
    case 16: hipLaunchKernelGGL((k_bwd_front_b<T, 256>), g, blk, 0, st, b); break;
    case 17: hipLaunchKernelGGL(k_rect_direct<T>, g, blk, 0, st, b); break;
    default: return hipErrorInvalidValue;
    }
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
template <typename T, int CH>
static void launch_cs_ch(int fwd, bool uni, dim3 g, const SolveArgs &b, int count, int64_t C, int nrhs, hipStream_t st)
{
    const dim3 blk(256);
    if (fwd) {
        if (uni) hipLaunchKernelGGL((k_cs_sweep<T, true, true, CH>), g, blk, 0, st, b, count, C, nrhs);
        else hipLaunchKernelGGL((k_cs_sweep<T, false, true, CH>), g, blk, 0, st, b, count, C, nrhs);
    } else {
        if (uni) hipLaunchKernelGGL((k_cs_sweep<T, true, false, CH>), g, blk, 0, st, b, count, C, nrhs);
        else hipLaunchKernelGGL((k_cs_sweep<T, false, false, CH>), g, blk, 0, st, b, count, C, nrhs);
    }
}
template <typename T>
hipError_t launch_cs_level(int fwd, const SolveArgs &a, const int32_t *lists, int list_off, int count, int64_t C,
                           int nrhs, int ch, hipStream_t st)
{
    if (count <= 0) return hipSuccess;
    SolveArgs b = a;
    b.list = lists + list_off;
    const int64_t items = (int64_t)count * ((C + 63) / 64);
    const dim3 g((unsigned)std::min<int64_t>((items + 3) / 4, 16384));
    const bool uni = nrhs % 64 == 0;
    switch (ch) {
    case 8: launch_cs_ch<T, 8>(fwd, uni, g, b, count, C, nrhs, st); break;
    case 16: launch_cs_ch<T, 16>(fwd, uni, g, b, count, C, nrhs, st); break;
    case 32: launch_cs_ch<T, 32>(fwd, uni, g, b, count, C, nrhs, st); break;
    case 64:   // FP64 stays at 32 rows: 64 would take 128 KB of gather rows per workgroup
        if constexpr (sizeof(T) == 4) launch_cs_ch<T, 64>(fwd, uni, g, b, count, C, nrhs, st);
        else launch_cs_ch<T, 32>(fwd, uni, g, b, count, C, nrhs, st);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
template <typename T>
hipError_t launch_cs_prep(const double *B, int64_t ldb, int64_t b_mat_stride, const double *rs, int64_t rs_stride,
                          const int32_t *qinv, T *w0, int64_t n, int nrhs, int64_t C, hipStream_t st)
{
    const int64_t gx = (C + 63) / 64, gy = std::max<int64_t>(1, std::min<int64_t>((n + 63) / 64, 65535));
    if (C + 64 > INT32_MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_cs_prep<T>, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, st, B, ldb, b_mat_stride, rs,
                       rs_stride, qinv, w0, n, nrhs, C);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_cs_finish(const T *x, const int32_t *qinv, double *X, int64_t ldx, int64_t x_mat_stride, int64_t n,
                            int nrhs, int64_t C, int accumulate, hipStream_t st)
{
    const int64_t gx = (C + 63) / 64, gy = std::max<int64_t>(1, std::min<int64_t>((n + 63) / 64, 65535));
    if (C + 64 > INT32_MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_cs_finish<T>, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, st, x, qinv, X, ldx,
                       x_mat_stride, n, nrhs, C, accumulate);
    return hipGetLastError();
}
hipError_t launch_residual(const int32_t *rp, const int32_t *ci, const int32_t *vmap, const double *ax,
                           int64_t ax_stride, const double *X, int64_t ldx, const double *B, int64_t ldb, double *R,
                           int64_t n, int64_t nrhs, int64_t msx, int64_t msb, int nmat, hipStream_t st)
{
    const int64_t blocks = (n + 255) / 256 * ((nrhs + RES_RC - 1) / RES_RC);
    if (blocks > INT32_MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_residual, dim3((unsigned)blocks, nmat), dim3(256), 0, st, rp, ci, vmap, ax, ax_stride, X, ldx,
                       B, ldb, R, n, nrhs, msx, msb);
    return hipGetLastError();
}
hipError_t launch_norms(const double *R, const double *X, const double *B, int64_t n, int64_t nrhs, double *part,
                        double *out, int nmat, hipStream_t st)
{
    const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(NORM_BLOCKS, (n + 2047) / 2048));
    hipLaunchKernelGGL(k_norms_part, dim3(nb, (unsigned)nrhs, nmat), dim3(256), 0, st, R, X, B, n, nrhs, part);
    const int64_t ncols = nrhs * nmat;
    hipLaunchKernelGGL(k_norms_fold, dim3((unsigned)((ncols + 63) / 64)), dim3(64), 0, st, part, ncols, nb, out);
    return hipGetLastError();
}
hipError_t launch_anorm(const int32_t *rp, const int32_t *vmap, const double *ax, int64_t ax_stride, int64_t n,
                        double *part, double *out, int nmat, hipStream_t st)
{
    const int nblk = 256;
    hipLaunchKernelGGL(k_anorm_part, dim3(nblk, nmat), dim3(256), 0, st, rp, vmap, ax, ax_stride, n, part);
    hipLaunchKernelGGL(k_anorm_fold, dim3((nmat + 63) / 64), dim3(64), 0, st, part, nblk, nmat, out);
    return hipGetLastError();
}
hipError_t launch_rowfac(const int32_t *prow, int32_t *rowfac, int64_t n, hipStream_t st)
{
    int blocks = (int)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_rowfac, dim3(blocks), dim3(256), 0, st, prow, rowfac, n);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_export(const SnDev *sn, int64_t nsuper, const T *store, const int32_t *upd, const int32_t *rowfac,
                         const int64_t *Lp, const int64_t *Up, const int64_t *uoff, int64_t *Li, double *Lx,
                         int64_t *Ui, double *Ux, hipStream_t st)
{
    hipLaunchKernelGGL(k_export<T>, dim3((unsigned)nsuper), dim3(256), 0, st, sn, store, upd, rowfac, Lp, Up, uoff, Li,
                       Lx, Ui, Ux);
    return hipGetLastError();
}

#define MF_INST(T)                                                                                                \
    template hipError_t launch_solve_level<T>(int, const SolveArgs &, const int32_t *, int, int, int, hipStream_t); \
    template hipError_t launch_solve_merged<T>(int, int, const SolveArgs &, const int32_t *, int, int, int, int, int, \
                                               hipStream_t);                                                        \
    template hipError_t launch_prep_rhs<T>(const double *, int64_t, const double *, int64_t, const int32_t *,   \
                                           T *, int64_t, int64_t, int64_t, int, hipStream_t);                   \
    template hipError_t launch_finish_x<T>(const T *, const int32_t *, double *, int64_t, int64_t, int64_t,     \
                                           int, int64_t, int, hipStream_t);                                     \
    template hipError_t launch_cs_level<T>(int, const SolveArgs &, const int32_t *, int, int, int64_t, int, int, \
                                           hipStream_t);                                                        \
    template hipError_t launch_cs_prep<T>(const double *, int64_t, int64_t, const double *, int64_t,            \
                                          const int32_t *, T *, int64_t, int, int64_t, hipStream_t);            \
    template hipError_t launch_cs_finish<T>(const T *, const int32_t *, double *, int64_t, int64_t, int64_t,    \
                                            int, int64_t, int, hipStream_t);                                    \
    template hipError_t launch_export<T>(const SnDev *, int64_t, const T *, const int32_t *, const int32_t *,   \
                                         const int64_t *, const int64_t *, const int64_t *, int64_t *, double *, \
                                         int64_t *, double *, hipStream_t);
MF_INST(double)
MF_INST(float)
#undef MF_INST

}  // namespace mf
