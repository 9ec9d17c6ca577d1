// mf_kernels.h -- device kernels of the multifrontal numeric phase (gfx950, wave64).
//
// Storage (see SnDev): a front lives in the workspace arena while its level is factored, as a
// dense nf x nf column-major block (ld nf) with rows/cols [pivot rows c1..c2 | update rows],
// the pre-pivoting order of src/multifrontal.jl:206-211.  When the level is done the panel
// [L11\U11; L21] (rows permuted by the in-block pivoting) and U12 go to the factor store, and
// the Schur complement F22 stays in the arena as the contribution block its parent reads,
// until the parent has assembled it.
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "mf_internal.h"

namespace mf {

constexpr int PANEL_KB = 16;   // panel width of the large-front getrf (pivoted schedule)
constexpr int SPEC_KB = 64;    // block width of the speculative (interchange-free) schedule
constexpr int SPEC_ROWS = 64;  // rows (or columns) per workgroup of the speculative trsm
// per front of a speculative diagonal launch: 1/u_jj and a flag for zero or tiny pivots (the
// trsm reads the factored block itself from the front)
constexpr int SPEC_SLOT = 2 * SPEC_KB;
constexpr int INV_LEAF = 64;   // diagonal block of the recursive pivot-block inversion
constexpr int INV_NP = 128;    // fronts with np > INV_NP keep inv(L11), inv(U11) for the solves
constexpr int PINFO_STRIDE = 1 + 4 * PANEL_KB;  // per-front panel permutation record
constexpr int GEMM_BM = 64, GEMM_BN = 64, GEMM_BK = 16;
// Tile shapes of the grouped MFMA gemm (rows x columns of C per 256-thread workgroup).  All
// shapes accumulate every C element over k in the same order (BK = 16 slices, MFMA k = 4),
// so the choice of tile never changes a result bit.
constexpr int GEMM_NTILES = 4;
constexpr int GEMM_TILE_BF16 = 64;   // the BF16 Schur-complement kernel (packed operands)
// BF16 Schur complements: operands packed in tiles of PK_ROWS rows (of A, or columns of B) x
// PK_K (k), each with its own power-of-two exponent; the gemm covers PK_ROWS x PK_ROWS of C
constexpr int PK_ROWS = 128, PK_K = 64, PK_TILE = PK_ROWS * PK_K;
constexpr int GEMM_TILE_M[GEMM_NTILES] = {64, 128, 64, 128};
constexpr int GEMM_TILE_N[GEMM_NTILES] = {64, 64, 128, 128};
// tile policy: size buckets of sqrt(m n) (upper bounds, the last open-ended) for the two gemm
// uses (0: panel updates, 1: Schur complements)
constexpr int GEMM_NBUCKETS = 6;
constexpr int GEMM_BUCKET_HI[GEMM_NBUCKETS - 1] = {96, 192, 384, 768, 1536};
inline int gemm_bucket(int64_t m, int64_t n)
{
    const double s = __builtin_sqrt((double)m * (double)n);
    int b = 0;
    while (b < GEMM_NBUCKETS - 1 && s >= GEMM_BUCKET_HI[b]) b++;
    return b;
}

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T> struct Tiny;
template <> struct Tiny<double> { static __device__ __forceinline__ double v() { return DBL_MIN; } };
template <> struct Tiny<float> { static __device__ __forceinline__ float v() { return FLT_MIN; } };

// Value of an original entry: (Rs .* A)[i, j] as apply_row_scaling forms it
// (src/scaling.jl:118-126): one FP64 product, then rounded to the front type.
template <typename T>
__device__ __forceinline__ T scaled_value(const double *ax, const double *rs, const int32_t *arow, int src)
{
    double v = ax[src];
    if (rs) v *= rs[arow[src]];
    return (T)v;
}

// Wave-level argmax of |v| with the first (smallest) row index winning ties, the rule of
// BLAS idamax that LAPACK getrf uses (src/multifrontal.jl:248 lu!(A11, RowMaximum())).
template <typename T>
__device__ __forceinline__ void wave_argmax(T &best, int &bi)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        T ob = __shfl_xor(best, off);
        int oi = __shfl_xor(bi, off);
        if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
}

// value held by lane l of the wave (l uniform), via v_readlane
__device__ __forceinline__ double lane_value(double v, int l)
{
    const unsigned long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(unsigned)u, l);
    const int hi = __builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
    return __longlong_as_double(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float lane_value(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// a value known to be equal in all lanes, moved to scalar registers
__device__ __forceinline__ double wave_uniform(double v)
{
    const unsigned long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(unsigned)u);
    const int hi = __builtin_amdgcn_readfirstlane((int)(unsigned)(u >> 32));
    return __longlong_as_double(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float wave_uniform(float v)
{
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

struct FactorArgs {
    const SnDev *sn;
    const int32_t *list;        // fronts of this launch (or columns as int2 for L_ASM_COLS / L_STORE)
    const int32_t *orig_row, *orig_col, *orig_src, *orig_colptr;
    const int32_t *child_list, *rmap;
    const int32_t *ext_colptr;
    const ExtRec *ext;
    const int32_t *arow;
    const double *ax;
    const double *rs;
    void *ws;                   // workspace arena (fronts during their level, contribution blocks)
    void *store;                // factor store (panels, U12, inverses)
    int32_t *prow;
    int32_t *info;
    int32_t *pinfo;             // per-front panel swap composition (PINFO_STRIDE ints each)
    int32_t *flag;              // per-matrix: set when the speculative schedule needs pivoting
    const int32_t *only;        // per-matrix: when set, matrices with only[mat] == 0 are skipped
    void *spec;                 // SPEC_SLOT elements per front of a speculative diagonal launch
    void *pack;                 // BF16 packed operand tiles (PK_TILE bf16 each)
    int32_t *pexp;              // their exponents
    int64_t ws_stride, store_stride, ax_stride, rs_stride, n_stride, pinfo_stride, spec_stride, pack_tiles;
};

// true when matrix `mat` of the launch is to be skipped (a redo of only the matrices whose
// interchange-free attempt failed)
__device__ __forceinline__ bool skip_matrix(const int32_t *only, int mat) { return only && only[mat] == 0; }

struct SolveArgs {
    const SnDev *sn;
    const int32_t *list;
    const int32_t *child_list, *rmap, *upd;
    const void *store;   // factor store (SnDev::foff, u12off, ioff)
    const int32_t *prow;
    void *w0;      // permuted, scaled RHS in V-row order (n x nrhs)
    void *y;       // forward result in factor-row order (n x nrhs)
    void *x;       // backward result in factor-column order (n x nrhs)
    void *uv;      // update vectors (nupd x nrhs)
    void *part;    // split-K partial sums of the rectangular products
    void *zbuf;    // large fronts: permuted pivot part (forward) / U12-corrected part (backward)
    const int32_t *ecolptr;  // extend-add lists of the factorization (per front position)
    const int32_t *eupd;     // their entries as offsets of the child's update rows in uv
    const int32_t *egpos;    // the front position (ecolptr index) of every entry
    int64_t n, nupd, nrhs;
    int64_t store_stride, n_stride, part_stride;
};

// ---- launchers (mf_factor_kernels.hip, mf_solve_kernels.hip) ----
template <typename T>
hipError_t launch_gemm_tile(int tile, int use, const GemmProb *probs, const GemmTile *tiles, int count, void *pool,
                            int64_t pool_stride, int nmat, const int32_t *only, hipStream_t st,
                            const void *pack = nullptr, const int32_t *pexp = nullptr, int64_t pack_tiles = 0);
hipError_t launch_copy(void *dst, const void *src, int64_t bytes, hipStream_t st);
template <typename T>
hipError_t launch_factor_step(const Launch &L, const FactorArgs &base, const int32_t *lists,
                              const GemmProb *probs, const GemmTile *tiles, int nmat, hipStream_t st);
int panel_variant_rows(int v);
// which: 0 forward triangular, 1/2 forward rectangular partials/sum (L21), 3/4 backward
// rectangular partials/sum (U12), 5 backward triangular, 6/7 inv(L11) partials/sum, 8/9
// inv(U11) partials/sum, 10 forward gather, 11-16 fused per-front sweeps, 17 forward L21
// product without split-K
template <typename T>
hipError_t launch_solve_level(int which, const SolveArgs &a, const int32_t *lists, int list_off, int count, int nmat,
                              hipStream_t st);
// one launch for a level's one-wave fronts (woff, wcount) and workgroup fronts (boff, bcount) of
// the fused sweeps; small: every workgroup front has nf <= 256
template <typename T>
hipError_t launch_solve_merged(int fwd, int small, const SolveArgs &a, const int32_t *lists, int boff, int bcount,
                               int woff, int wcount, int nmat, hipStream_t st);
template <typename T>
hipError_t launch_prep_rhs(const double *B, int64_t ldb, const double *rs, int64_t rs_stride, const int32_t *qf,
                           T *w0, int64_t n, int64_t nrhs, int64_t b_mat_stride, int nmat, hipStream_t st);
template <typename T>
hipError_t launch_finish_x(const T *x, const int32_t *qf, double *X, int64_t ldx, int64_t n, int64_t nrhs,
                           int accumulate, int64_t x_mat_stride, int nmat, hipStream_t st);
// Column-parallel sweeps (one lane per right-hand-side column of the batch, vectors
// interleaved by column: entry i of column g = mat * nrhs + r at i * C + g, C = nmat * nrhs)
template <typename T>
hipError_t launch_cs_level(int fwd, const SolveArgs &a, const int32_t *lists, int list_off, int count, int64_t C,
                           int nrhs, int ch, hipStream_t st);
template <typename T>
hipError_t launch_cs_prep(const double *B, int64_t ldb, int64_t b_mat_stride, const double *rs, int64_t rs_stride,
                          const int32_t *qinv, T *w0, int64_t n, int nrhs, int64_t C, hipStream_t st);
template <typename T>
hipError_t launch_cs_finish(const T *x, const int32_t *qinv, double *X, int64_t ldx, int64_t x_mat_stride, int64_t n,
                            int nrhs, int64_t C, int accumulate, hipStream_t st);
hipError_t launch_residual(const int32_t *rp, const int32_t *ci, const int32_t *vmap, const double *ax,
                           int64_t ax_stride, const double *X, int64_t ldx, const double *B, int64_t ldb, double *R,
                           int64_t n, int64_t nrhs, int64_t msx, int64_t msb, int nmat, hipStream_t st);
hipError_t launch_norms(const double *R, const double *X, const double *B, int64_t n, int64_t nrhs, double *part,
                        double *out, int nmat, hipStream_t st);
hipError_t launch_anorm(const int32_t *rp, const int32_t *vmap, const double *ax, int64_t ax_stride, int64_t n,
                        double *part, double *out, int nmat, hipStream_t st);
hipError_t launch_rowfac(const int32_t *prow, int32_t *rowfac, int64_t n, hipStream_t st);
template <typename T>
hipError_t launch_export(const SnDev *sn, int64_t nsuper, const T *store, const int32_t *upd, const int32_t *rowfac,
                         const int64_t *Lp, const int64_t *Up, const int64_t *uoff, int64_t *Li, double *Lx,
                         int64_t *Ui, double *Ux, hipStream_t st);
hipError_t launch_fill_i32(int32_t *p, int32_t v, int64_t n, hipStream_t st);
// p[i] = v where mask[i] != 0 (mask NULL: everywhere)
hipError_t launch_fill_i32_masked(int32_t *p, int32_t v, const int32_t *mask, int64_t n, hipStream_t st);

}  // namespace mf
