// mf_internal.h -- data structures shared by the host plan builder and the HIP kernels.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace mf {

using i64 = int64_t;

// Result of the host symbolic analysis (0-based), the arrays of SymbolicMF
// (src/symbolic.jl:203-209) with colstruct kept only for each supernode's last column.
struct Symbolic {
    i64 n = 0;
    std::vector<i64> qf, parent, sstart;
    std::vector<std::vector<i64>> upd;       // update rows of each supernode (ascending)
    std::vector<std::vector<i64>> children;  // child supernodes (ascending)
    i64 ncolstruct = 0;                      // sum over columns of |colstruct[j]|
};

bool analyze(i64 n, const i64 *Ap, const i64 *Ai, int ordering, const i64 *q_in, Symbolic &out,
             std::string &err);
// relaxed supernode amalgamation of an analysis (in place); returns the merges made
int64_t relax_supernodes(Symbolic &S, double zmax, int64_t npmax);

// Per-supernode record in device memory.  Two storage areas per matrix:
//   * the factor store keeps what the solves and the export read for the life of the plan:
//     the panel [L11\U11; L21] (nf x np, column-major, ld nf) at foff and U12 (np x nu,
//     column-major, ld np) at u12off, plus the inverses of large pivot blocks at ioff;
//   * the workspace arena holds a front while its level is factored (nf x nf, ld nf, at woff)
//     and its contribution block F22 until the parent has assembled it (at cboff, leading
//     dimension cbld).  Arena offsets come from a liveness plan over the etree levels, so
//     fronts whose lifetimes do not overlap share memory (src/multifrontal.jl:180-187,319-332
//     keeps a LIFO arena for the same reason).
struct SnDev {
    i64 foff;          // factor store: panel [L11\U11; L21]
    i64 u12off;        // factor store: U12
    i64 woff;          // arena: whole front during its level; -1: the fused kernel keeps it in LDS
    i64 cboff;         // arena: contribution block
    i64 ioff;          // >= 0: inv(L11) at ioff and inv(U11) at ioff + np*np (ld np) in the store
    int32_t c1, np, nu, nf;
    int32_t upd_off;   // offset of this supernode's rows in upd[] / rmap[] / uoff[]
    int32_t orig_off;  // offset of this front's original entries (sorted by column)
    int32_t norig;
    int32_t child_off; // offset in child list (children stored in REVERSE order)
    int32_t nchild;
    int32_t fcol;      // index of this front's first column in the per-front-column arrays
    int32_t cbld;      // leading dimension of the contribution block (nf in a whole front, nu alone)
    int32_t pad;
};

// Grouped gemm problem, column-major, offsets in elements into one matrix's pool:
//   C -= A B (flags 0), C += A B (GEMM_ADD), C = -A B (GEMM_SET), C = A B (GEMM_SET|GEMM_ADD).
// A triangular operand (A_LOWER: A[i][k] = 0 for k > i, A_UPPER: for k < i, B_LOWER: B[k][j] = 0
// for k < j, B_UPPER: for k > j, zeros stored) lets a tile skip the k-slices it knows are zero.
enum GemmFlags : int32_t {
    GEMM_ADD = 1, GEMM_SET = 2, GEMM_A_LOWER = 4, GEMM_A_UPPER = 8, GEMM_B_LOWER = 16, GEMM_B_UPPER = 32
};
struct GemmProb {
    i64 aoff, boff, coff;
    int32_t m, n, k, lda, ldb, ldc;
    int32_t flags;
    int32_t pad;
    i64 apk, bpk;   // BF16 Schur complements: first packed tile of A and of B (tile index in the pack buffer)
};

// One child contribution to one parent front column (extend-add): the child's contribution-block
// column at src (arena offset), its nu rows mapped into the parent by rmap[roff .. roff + nu).
struct ExtRec {
    i64 src;
    int32_t roff, nu;
};

// Grouped gemm tile: problem index and tile origin.
struct GemmTile {
    int32_t prob, tm, tn, pad;
};

enum LaunchKind : int32_t {
    L_FUSED_SMALL = 0,  // one workgroup per small front: assemble + factor in LDS
    L_ASM_COLS = 1,     // one workgroup per large-front column: assemble in LDS
    L_PANEL = 2,        // one workgroup per large front: factor a panel of the pivot block
    L_GEMM = 3,         // grouped C -= A B (panel updates and the Schur update)
    L_PANEL_APPLY = 4,  // row swaps + U-row solve + L21 solve after a panel (many workgroups)
    L_SPEC_DIAG = 5,    // speculative schedule: factor the diagonal block without interchanges
    L_SPEC_TRSM = 6,    // speculative schedule: L rows / U columns of the block + pivot check
    L_INV_LEAF = 7,     // inverses of the 64x64 diagonal blocks of L11 and U11 (large fronts)
    L_SPEC_DIAG_GEMM = 8,  // next diagonal block factorization alongside the bulk of the trailing update
    L_STORE = 9,        // copy the factored panel and U12 of a level's large fronts to the store
    L_PACK_BF16 = 10,   // BF16 Schur complements: scale, round and pack the operand tiles
};

struct Launch {
    int32_t kind;
    int32_t count;      // workgroups (fronts / columns / tiles)
    int32_t list_off;   // offset into the int32 work-list array (fronts or columns)
    int32_t k;          // panel start column (L_PANEL)
    int32_t variant;    // kernel variant (panel rows-per-thread / LDS size selector)
    int32_t prob_off;   // L_GEMM: offset into problems; tiles at list_off in tile array
    int32_t shmem;      // dynamic LDS bytes
    int32_t level;
    int32_t tile;       // L_GEMM: tile shape (GEMM_TILE_* index)
    int32_t on_store;   // L_GEMM: operands in the factor store (pivot-block inversion), else the arena
    int32_t tile_off;   // L_SPEC_DIAG_GEMM: offset of its gemm tiles (its fronts are at list_off)
    int32_t ntiles;     // L_SPEC_DIAG_GEMM: gemm tiles before the diagonal-block workgroups
    int32_t dprob;      // L_SPEC_DIAG_GEMM: problem (from prob_off) of the first diagonal block's
                        // own panel update, one per diagonal workgroup; -1: none
    float tau;          // speculative schedule: multiplier bound 1/tau
    double flops;       // algorithmic flops of the launch (host-side bookkeeping)
    double bytes;       // algorithmic HBM bytes of the launch
};

}  // namespace mf
