// mf_plan.hip -- plan construction (the static device schedule) and the C ABI of mfgpu.h.
//
// The symbolic structure fixes every size of the numeric phase: front shapes, where each
// original entry and each child column lands, the L/U CSC offsets, the etree levels.  The
// plan turns it, once per pattern, into device arrays and a flat list of kernel launches
// (level by level: small fronts fused in LDS, large fronts as column assembly -> panel
// getrf -> grouped gemm updates -> Schur gemm).  mf_factor only replays that list.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <new>
#include <string>
#include <vector>

#include "../../include/mfgpu.h"
#include "mf_internal.h"
#include "mf_kernels.h"

using namespace mf;

static constexpr int64_t SOLVE_CK_HOST = 64;    // column chunk of the split-K solve products (k_rect_part)
static constexpr int64_t SOLVE_FUSE_MIN = 128;  // fewer mid-size fronts in a level: split pipeline
static constexpr int64_t FRONT_MAXNF_HOST = 1024;  // largest front of the fused per-front sweeps
static constexpr int64_t CS_CH_HOST = 64;         // register chunk of the column-parallel sweeps
// MF_SOLVE_AUTO takes the column-parallel sweeps from this many right-hand-side columns (the
// batch's matrices times their right-hand sides) when no pivot block is larger than
// CS_MAXNP: each lane runs a front's substitution by itself, so a wide pivot block is a long
// sequential chain
static constexpr int64_t CS_MIN_COLS = 512;
static constexpr int64_t CS_MAXNP = 1024;

static thread_local std::string g_err;
// gemm tile policy per precision (0: FP64, 1: FP32/BF16), use (panel update / Schur) and size
// bucket; read when a plan is created.  Default: the 64 x 64 tile everywhere.
static int32_t g_tile_policy[2][2 * GEMM_NBUCKETS] = {};

// Workgroup i of a launch runs on XCD i % 8 (round-robin dispatch over the 8 XCDs, each with
// its own L2).  Reorder a launch's tiles [begin, end) so that XCD x receives the contiguous
// x-th eighth of the original (problem, column block, row block) order: neighbouring tiles,
// which share rows of A or columns of B, then share an L2.
static void xcd_order(std::vector<GemmTile> &t, size_t begin)
{
    constexpr size_t NX = 8;
    const size_t cnt = t.size() - begin, per = cnt / NX;
    if (per < 2) return;
    std::vector<GemmTile> o(t.begin() + begin, t.end());
    for (size_t p = 0; p < per * NX; p++) t[begin + p] = o[(p % NX) * per + p / NX];
}

// Launches whose problems differ in k (the Schur complements of a level's fronts, the pivot-block
// inversions) would put all of one front's heavy tiles on one or two XCDs under xcd_order.  Here
// every XCD receives a contiguous eighth of every problem's tiles (so each XCD carries an eighth
// of each problem's work, and its tiles of one problem still share rows of A or columns of B),
// and runs the problems in order of decreasing k, heaviest first.  Tiles [begin, end) must be
// grouped by problem; kw[prob] is the k of problem `prob` (the index stored in the tile).  A
// problem with bcyc[prob] set (a triangular B: a tile's k range grows or shrinks with its column
// block) has its column blocks dealt out cyclically, so that each eighth holds column blocks
// from the whole width.
static void xcd_order_by_problem(std::vector<GemmTile> &t, size_t begin, const std::vector<int64_t> &kw,
                                 const std::vector<char> &bcyc = {})
{
    constexpr size_t NX = 8;
    const size_t cnt = t.size() - begin;
    if (cnt == 0) return;
    std::vector<std::pair<size_t, size_t>> runs;   // [first, last) of each problem's tiles
    for (size_t i = begin; i < t.size(); i++)
        if (i == begin || t[i].prob != t[i - 1].prob) runs.push_back({i, i + 1});
        else runs.back().second = i + 1;
    for (const auto &r : runs) {
        if ((size_t)t[r.first].prob >= bcyc.size() || !bcyc[t[r.first].prob]) continue;
        std::vector<int32_t> cols;   // distinct column-block origins, ascending
        for (size_t i = r.first; i < r.second; i++) cols.push_back(t[i].tn);
        std::sort(cols.begin(), cols.end());
        cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
        auto key = [&](const GemmTile &x) {
            const int64_t cb = std::lower_bound(cols.begin(), cols.end(), x.tn) - cols.begin();
            return std::make_tuple(cb % (int64_t)NX, cb / (int64_t)NX, x.tm);
        };
        std::stable_sort(t.begin() + r.first, t.begin() + r.second,
                         [&](const GemmTile &a, const GemmTile &b) { return key(a) < key(b); });
    }
    std::stable_sort(runs.begin(), runs.end(), [&](const std::pair<size_t, size_t> &a, const std::pair<size_t, size_t> &b) {
        return kw[t[a.first].prob] > kw[t[b.first].prob];
    });
    std::vector<GemmTile> q[NX];
    size_t c = 0;   // next XCD to receive an extra tile: keeps the queue lengths within one
    for (const auto &r : runs) {
        const size_t T = r.second - r.first, base = T / NX, rem = T % NX;
        size_t pos = r.first;
        for (size_t j = 0; j < NX; j++) {
            const size_t take = base + (j < rem ? 1 : 0);
            std::vector<GemmTile> &dst = q[(c + j) % NX];
            dst.insert(dst.end(), t.begin() + pos, t.begin() + pos + take);
            pos += take;
        }
        c = (c + rem) % NX;
    }
    // workgroup i runs on XCD i % 8: queue x has exactly ceil((cnt - x) / 8) tiles
    std::vector<GemmTile> o(cnt);
    for (size_t i = 0; i < cnt; i++) o[i] = q[i % NX][i / NX];
    std::copy(o.begin(), o.end(), t.begin() + begin);
}

static int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            return fail(e_ == hipErrorOutOfMemory ? MF_OOM : MF_HIP,                              \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                       \
        }                                                                                         \
    } while (0)

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t b)
    {
        if (p && bytes >= b) return hipSuccess;
        if (p) { (void)hipFree(p); p = nullptr; bytes = 0; }
        hipError_t e = hipMalloc(&p, b > 0 ? b : 16);
        if (e == hipSuccess) bytes = b;
        return e;
    }
    template <typename T> T *as() const { return reinterpret_cast<T *>(p); }
};

template <typename T>
static hipError_t upload(DevBuf &b, const std::vector<T> &v)
{
    hipError_t e = b.alloc(v.size() * sizeof(T));
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
}

struct mf_symbolic {
    Symbolic S;
};

// Buffers of one solve call: the sweep vectors in the front precision (w0 permuted scaled
// RHS, y forward result, x backward result, uv update vectors, part split-K partials, zbuf
// large-front staging) and the FP64 refinement vectors (B, X, R) with the norm partials.  They
// are carved from one region: a plan-owned allocation, or the file-scope scratch array of the
// allocation-free entry point.
struct SolveBufs {
    void *w0 = nullptr, *y = nullptr, *x = nullptr, *uv = nullptr, *part = nullptr, *zbuf = nullptr;
    double *B = nullptr, *X = nullptr, *R = nullptr, *norms = nullptr, *npart = nullptr;
};

struct mf_plan {
    // options
    int prec = MF_FP64;
    int small_nf = 64;
    int use_graph = 0;
    int nbatch = 1;
    // sizes
    int64_t n = 0, nnz = 0, nsuper = 0, nupd = 0, nlevels = 0, maxnf = 0;
    int64_t store_elems = 0;   // factor store per matrix: panels | U12 | inverses | inversion temps
    int64_t ws_elems = 0;      // workspace arena per matrix (liveness-planned peak)
    int64_t nnzL = 0, nnzU = 0;
    double flops = 0, gemm_flops = 0, gemm_flops_big = 0;
    // host copies
    std::vector<int64_t> qf, Lp, Up, uoff;
    std::vector<SnDev> sn;
    std::vector<Launch> sched[2];   // 0: partial-pivoting panels, 1: interchange-free blocks
    bool has_large = false;
    int pivoting = MF_PIVOT_AUTO;
    float pivot_tol = 0.f;
    int last_sched = 0;
    int last_redone = 0;            // matrices the last call factored on the pivoted schedule
    std::vector<int32_t> lvl_off, lvl_cnt;  // solve lists per level
    // per level: offsets/counts of the forward/backward split-K part and sum item lists
    std::vector<int32_t> lvl_fpo, lvl_fpc, lvl_fso, lvl_fsc, lvl_bpo, lvl_bpc, lvl_bso, lvl_bsc;
    std::vector<int32_t> lvl_ilpo, lvl_ilpc, lvl_ilso, lvl_ilsc, lvl_iupo, lvl_iupc, lvl_iuso, lvl_iusc;
    std::vector<int32_t> lvl_ntri;   // large fronts of the level without stored inverses
    std::vector<int32_t> lvl_fdo, lvl_fdc;   // forward L21 product without split-K (k_rect_direct)
    std::vector<int32_t> lvl_go, lvl_gc;   // forward gather items {front, pos0}
    // fronts swept by the fused per-front kernels (no stored inverse): nf <= 64 and
    // 64 < nf <= FRONT_MAXNF; the level lists above hold only the other fronts
    std::vector<int32_t> lvl_s64o, lvl_s64c, lvl_s1ko, lvl_s1kc;
    std::vector<int32_t> lvl_s1k_small;   // every such front of the level has nf <= 256
    std::vector<int32_t> lvl_allo, lvl_allc;   // every front of the level (column-parallel sweeps)
    std::vector<int32_t> lvl_csch;             // their register chunk: 75th percentile of nf, 8..64
    int64_t maxnp = 0;                         // largest pivot block
    int solve_path = MF_SOLVE_AUTO;
    int64_t factor_elems = 0, inv_elems = 0, tmp_elems = 0;  // store = factors | inverses | gemm temps
    int64_t ws_sum_elems = 0;  // what every front kept resident would take (for the stats)
    int64_t part_slots = 0;  // partial-sum slots (256 rows x nrhs each) the solve needs
    std::vector<double> lvl_lflops, lvl_uflops, lvl_lbytes, lvl_ubytes, lvl_vec;
    // device
    DevBuf d_sn, d_upd, d_rmap, d_child, d_orow, d_ocol, d_osrc, d_ocolptr, d_ecolptr, d_ext, d_eupd, d_arow;
    DevBuf d_lists, d_probs, d_tiles, d_solvelist, d_qf, d_qinv, d_egpos;
    DevBuf d_csr_rp, d_csr_ci, d_csr_vmap;
    DevBuf d_store, d_ws, d_prow, d_info, d_ax, d_rs, d_pinfo, d_flag, d_spec, d_redo;
    int64_t spec_slots = 0;         // most fronts in one speculative diagonal launch
    int64_t pack_tiles = 0;         // BF16 operand tiles one launch packs at most
    DevBuf d_pack, d_pexp;
    DevBuf d_solve, d_anorm, d_npart;   // solve buffers (carved as SolveBufs), ||A||, its partials
    SolveBufs sb;                       // the buffers the current solve call uses
    bool sb_scratch = false;            // sb points into the allocation-free scratch
    struct SolveGraph {
        hipGraphExec_t exec = nullptr;
        int64_t nrhs = 0;
        int nmat = 0;
        const double *rhs = nullptr;
        const void *w0 = nullptr;
    } sgraph[3];   // captured sweeps: [0] solve from B, [1] correction from R, [2] [0] with the top inversions
    DevBuf d_Lp, d_Up, d_uoff, d_Li, d_Lx, d_Ui, d_Ux, d_rowfac;
    int64_t solve_nrhs = 0;
    bool factored = false;
    int nmat_factored = 0;
    int rs_given = 0;
    hipStream_t stream = nullptr;
    // pivot-block inversions of the top level's fronts: run by the first solve after a
    // factorization on stream `side`, beside the sweeps of the levels below (which do not
    // read them), joined before the top level's sweep
    std::vector<Launch> inv_top;
    bool inv_pending = false;   // factored, top-level inverses not computed yet
    bool side_busy = false;     // inversions enqueued on `side` that no host sync has covered
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_inv = nullptr;
    hipGraphExec_t graph[3] = {nullptr, nullptr, nullptr};
    int graph_nmat[3] = {0, 0, 0};
    int graph_rs[3] = {0, 0, 0};
    // profiling (mf_plan_set_profiling)
    int prof = 0;
    double prof_ms[MF_PROF_CLASSES] = {0}, prof_flops[MF_PROF_CLASSES] = {0}, prof_bytes[MF_PROF_CLASSES] = {0};
    int64_t prof_n[MF_PROF_CLASSES] = {0};
    std::vector<hipEvent_t> ev;
    struct EvRec { int cls; size_t a, b; double flops, bytes; };
    std::vector<EvRec> evrec;
    size_t evused = 0;

    ~mf_plan()
    {
        if (side) (void)hipStreamSynchronize(side);
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_inv) (void)hipEventDestroy(ev_inv);
        if (side) (void)hipStreamDestroy(side);
        for (int g = 0; g < 3; g++)
            if (graph[g]) (void)hipGraphExecDestroy(graph[g]);
        for (int g = 0; g < 3; g++)
            if (sgraph[g].exec) (void)hipGraphExecDestroy(sgraph[g].exec);
        if (stream) (void)hipStreamDestroy(stream);
    }
    size_t esize() const { return prec == MF_FP64 ? 8 : 4; }

    hipError_t ev_mark(size_t &idx)
    {
        if (evused == ev.size()) {
            hipEvent_t e;
            hipError_t r = hipEventCreate(&e);
            if (r != hipSuccess) return r;
            ev.push_back(e);
        }
        idx = evused++;
        return hipEventRecord(ev[idx], stream);
    }
    // collect the recorded pairs after the stream has been synchronised
    hipError_t ev_collect()
    {
        for (const EvRec &r : evrec) {
            float ms = 0;
            hipError_t e = hipEventElapsedTime(&ms, ev[r.a], ev[r.b]);
            if (e != hipSuccess) return e;
            prof_ms[r.cls] += ms;
            prof_n[r.cls] += 1;
            prof_flops[r.cls] += r.flops;
            prof_bytes[r.cls] += r.bytes;
        }
        evrec.clear();
        evused = 0;
        return hipSuccess;
    }
};

// Runs `launch` between two profiling events when profiling is on.
template <typename F>
static hipError_t profiled(mf_plan &P, int cls, double flops, double bytes, F &&launch)
{
    if (!P.prof) return launch();
    size_t a, b;
    hipError_t e = P.ev_mark(a);
    if (e != hipSuccess) return e;
    e = launch();
    if (e != hipSuccess) return e;
    e = P.ev_mark(b);
    if (e != hipSuccess) return e;
    P.evrec.push_back({cls, a, b, flops, bytes});
    return hipSuccess;
}

static int launch_class(const Launch &L)
{
    switch (L.kind) {
    case L_FUSED_SMALL: return 0;
    case L_ASM_COLS: return 1;
    case L_PANEL: return 2;
    case L_PANEL_APPLY: return 2;
    case L_SPEC_DIAG: return 2;
    case L_SPEC_TRSM: return 2;
    case L_SPEC_DIAG_GEMM: return 2;
    case L_STORE: return 1;
    case L_PACK_BF16: return 1;
    default: return L.variant == 3 ? 7 : L.variant == 1 ? 4 : 3;
    }
}

extern "C" const char *mf_last_error(void) { return g_err.c_str(); }

// ------------------------------------------------------------------ symbolic API
extern "C" int mf_analyze(int64_t n, const int64_t *Ap, const int64_t *Ai, int32_t ordering, const int64_t *q,
                          mf_symbolic **out)
{
    if (!out) return fail(MF_BADARG, "out is NULL");
    *out = nullptr;
    if (n < 0 || !Ap || (n > 0 && !Ai)) return fail(MF_BADARG, "bad pattern arguments");
    if (ordering < 0 || ordering > 2) return fail(MF_BADARG, "unknown ordering");
    for (int64_t j = 0; j < n; j++)
        if (Ap[j + 1] < Ap[j]) return fail(MF_BADARG, "column pointers decrease");
    for (int64_t k = 0; k < Ap[n] - 1; k++)
        if (Ai[k] < 1 || Ai[k] > n) return fail(MF_BADARG, "row index out of range");
    std::unique_ptr<mf_symbolic> s(new (std::nothrow) mf_symbolic);
    if (!s) return fail(MF_OOM, "host allocation failed");
    std::string err;
    try {
        if (!analyze(n, Ap, Ai, ordering, q, s->S, err)) return fail(MF_BADARG, err);
    } catch (const std::bad_alloc &) {
        return fail(MF_OOM, "host allocation failed in analysis");
    }
    *out = s.release();
    return MF_OK;
}

extern "C" int mf_symbolic_sizes(const mf_symbolic *S, int64_t *nsuper, int64_t *nupd, int64_t *nchild,
                                 int64_t *ncolstruct)
{
    if (!S) return fail(MF_BADARG, "NULL symbolic");
    int64_t ns = (int64_t)S->S.sstart.size() - 1, nu = 0, nc = 0;
    for (auto &u : S->S.upd) nu += (int64_t)u.size();
    for (auto &c : S->S.children) nc += (int64_t)c.size();
    if (nsuper) *nsuper = ns;
    if (nupd) *nupd = nu;
    if (nchild) *nchild = nc;
    if (ncolstruct) *ncolstruct = S->S.ncolstruct;
    return MF_OK;
}

extern "C" int mf_symbolic_export(const mf_symbolic *S, int64_t *qf, int64_t *parent, int64_t *sstart,
                                  int64_t *upd_ptr, int64_t *upd_idx, int64_t *child_ptr, int64_t *child_idx)
{
    if (!S) return fail(MF_BADARG, "NULL symbolic");
    const Symbolic &s = S->S;
    int64_t n = s.n, ns = (int64_t)s.sstart.size() - 1;
    if (qf)
        for (int64_t j = 0; j < n; j++) qf[j] = s.qf[j] + 1;
    if (parent)
        for (int64_t j = 0; j < n; j++) parent[j] = s.parent[j] + 1;
    if (sstart)
        for (int64_t k = 0; k <= ns; k++) sstart[k] = s.sstart[k] + 1;
    int64_t o = 0;
    for (int64_t k = 0; k < ns; k++) {
        if (upd_ptr) upd_ptr[k] = o + 1;
        for (int64_t v : s.upd[k]) {
            if (upd_idx) upd_idx[o] = v + 1;
            o++;
        }
    }
    if (upd_ptr) upd_ptr[ns] = o + 1;
    o = 0;
    for (int64_t k = 0; k < ns; k++) {
        if (child_ptr) child_ptr[k] = o + 1;
        for (int64_t v : s.children[k]) {
            if (child_idx) child_idx[o] = v + 1;
            o++;
        }
    }
    if (child_ptr) child_ptr[ns] = o + 1;
    return MF_OK;
}

extern "C" int mf_symbolic_relax(mf_symbolic *S, double max_zero_fraction, int64_t max_small_np, int64_t *merged)
{
    if (!S) return fail(MF_BADARG, "NULL symbolic");
    if (!(max_zero_fraction >= 0.0 && max_zero_fraction < 1.0) || max_small_np < 0)
        return fail(MF_BADARG, "relaxation limits out of range");
    int64_t m = 0;
    try {
        m = relax_supernodes(S->S, max_zero_fraction, max_small_np);
    } catch (const std::bad_alloc &) {
        return fail(MF_OOM, "host allocation failed in the amalgamation");
    }
    if (merged) *merged = m;
    return MF_OK;
}

extern "C" void mf_symbolic_destroy(mf_symbolic *S) { delete S; }

// ------------------------------------------------------------------ plan
static int choose_panel_variant(int rows)
{
    for (int v = 0; v < 6; v++)
        if (panel_variant_rows(v) >= rows) return v;
    return -1;
}

static int build_plan(mf_plan &P, int64_t n, const int64_t *qf1, int64_t nsuper, const int64_t *sstart1,
                      const int64_t *upd_ptr1, const int64_t *upd_idx1, const int64_t *child_ptr1,
                      const int64_t *child_idx1, const int64_t *Ap1, const int64_t *Ai1, bool device)
{
    P.n = n;
    P.nsuper = nsuper;
    P.nnz = Ap1[n] - 1;
    if (n >= INT32_MAX / 2 || P.nnz >= INT32_MAX) return fail(MF_BADARG, "matrix too large for int32 indexing");
    // ---- validate and convert
    P.qf.resize(n);
    {
        std::vector<char> seen(n, 0);
        for (int64_t j = 0; j < n; j++) {
            int64_t v = qf1[j] - 1;
            if (v < 0 || v >= n || seen[v]) return fail(MF_BADARG, "qf is not a permutation");
            seen[v] = 1;
            P.qf[j] = v;
        }
    }
    if (sstart1[0] != 1 || sstart1[nsuper] != n + 1) return fail(MF_BADARG, "sstart must span 1..n+1");
    std::vector<int64_t> c1(nsuper), np(nsuper), nu(nsuper), nf(nsuper), uoff0(nsuper + 1, 0);
    std::vector<int64_t> super_of(n);
    for (int64_t s = 0; s < nsuper; s++) {
        c1[s] = sstart1[s] - 1;
        np[s] = sstart1[s + 1] - sstart1[s];
        if (np[s] <= 0) return fail(MF_BADARG, "empty supernode");
        nu[s] = upd_ptr1[s + 1] - upd_ptr1[s];
        if (nu[s] < 0) return fail(MF_BADARG, "bad upd_ptr");
        nf[s] = np[s] + nu[s];
        uoff0[s + 1] = uoff0[s] + nu[s];
        for (int64_t j = c1[s]; j < c1[s] + np[s]; j++) super_of[j] = s;
    }
    P.nupd = uoff0[nsuper];
    std::vector<int32_t> upd(P.nupd);
    for (int64_t s = 0; s < nsuper; s++) {
        int64_t last = c1[s] + np[s] - 1;
        for (int64_t t = 0; t < nu[s]; t++) {
            int64_t v = upd_idx1[upd_ptr1[s] - 1 + t] - 1;
            if (v <= last || v >= n || (t > 0 && v <= upd[uoff0[s] + t - 1]))
                return fail(MF_BADARG, "update rows must be ascending and beyond the supernode");
            upd[uoff0[s] + t] = (int32_t)v;
        }
    }
    std::vector<int64_t> parent_sn(nsuper, -1);
    std::vector<std::vector<int64_t>> kids(nsuper);
    for (int64_t s = 0; s < nsuper; s++)
        for (int64_t q = child_ptr1[s] - 1; q < child_ptr1[s + 1] - 1; q++) {
            int64_t c = child_idx1[q] - 1;
            if (c < 0 || c >= s || parent_sn[c] != -1) return fail(MF_BADARG, "bad child list");
            parent_sn[c] = s;
            kids[s].push_back(c);
        }
    // levels
    std::vector<int64_t> level(nsuper, 0);
    int64_t nlev = 0;
    for (int64_t s = 0; s < nsuper; s++) {
        for (int64_t c : kids[s]) level[s] = std::max(level[s], level[c] + 1);
        nlev = std::max(nlev, level[s] + 1);
    }
    P.nlevels = nlev;
    // rmap: position of each child update row in the parent front
    std::vector<int32_t> loc(n, -1), rmap(P.nupd);
    for (int64_t s = 0; s < nsuper; s++) {
        for (int64_t a = 0; a < np[s]; a++) loc[c1[s] + a] = (int32_t)a;
        for (int64_t t = 0; t < nu[s]; t++) loc[upd[uoff0[s] + t]] = (int32_t)(np[s] + t);
        for (int64_t c : kids[s])
            for (int64_t t = 0; t < nu[c]; t++) {
                int32_t l = loc[upd[uoff0[c] + t]];
                if (l < 0) return fail(MF_BADARG, "child update row missing from parent front");
                rmap[uoff0[c] + t] = l;
            }
        for (int64_t a = 0; a < np[s]; a++) loc[c1[s] + a] = -1;
        for (int64_t t = 0; t < nu[s]; t++) loc[upd[uoff0[s] + t]] = -1;
    }
    // factor store: per supernode the panel [L11\U11; L21] (nf x np) then U12 (np x nu)
    std::vector<int64_t> foff(nsuper), u12off(nsuper), fcol(nsuper + 1, 0);
    int64_t pe = 0;
    for (int64_t s = 0; s < nsuper; s++) {
        foff[s] = pe;
        pe += nf[s] * np[s];
        u12off[s] = pe;
        pe += np[s] * nu[s];
        fcol[s + 1] = fcol[s] + nf[s];
        P.maxnf = std::max(P.maxnf, nf[s]);
    }
    P.factor_elems = pe;
    // large fronts keep inv(L11) and inv(U11) (np x np each) after the factors
    std::vector<int64_t> ioff(nsuper, -1);
    for (int64_t s = 0; s < nsuper; s++)
        if (np[s] > INV_NP) {
            ioff[s] = pe;
            pe += 2 * np[s] * np[s];
        }
    P.inv_elems = pe - P.factor_elems;
    if (fcol[nsuper] >= INT32_MAX) return fail(MF_BADARG, "too many front columns");
    // workspace arena: a large front (nf > small_nf) holds nf x nf from its level through its
    // parent's level (the parent assembles its F22 from there); a small front, factored in LDS,
    // only its nu x nu contribution block.  Offsets come from a level-order simulation of a
    // best-fit allocator: blocks freed after level L are reused from level L + 1 on.
    std::vector<int64_t> woff(nsuper, -1), cboff(nsuper, 0), cbld(nsuper, 1);
    {
        std::vector<int64_t> sn_parent(nsuper, -1);
        for (int64_t s = 0; s < nsuper; s++)
            for (int64_t c : kids[s]) sn_parent[c] = s;
        std::vector<std::vector<int64_t>> starts(nlev), ends(nlev);
        std::vector<int64_t> need(nsuper, 0);
        for (int64_t s = 0; s < nsuper; s++) {
            const bool large = nf[s] > P.small_nf;
            need[s] = large ? nf[s] * nf[s] : nu[s] * nu[s];
            P.ws_sum_elems += nf[s] * nf[s];
            if (need[s] == 0) continue;
            need[s] = (need[s] + 15) / 16 * 16;   // 64-byte granules (FP32)
            starts[level[s]].push_back(s);
            ends[sn_parent[s] >= 0 ? level[sn_parent[s]] : level[s]].push_back(s);
        }
        std::multimap<int64_t, int64_t> by_size;     // free blocks: size -> offset
        std::map<int64_t, int64_t> by_off;           // free blocks: offset -> size
        int64_t top = 0;
        auto release = [&](int64_t off, int64_t sz) {
            auto nx = by_off.lower_bound(off);
            if (nx != by_off.end() && off + sz == nx->first) {   // merge with the next block
                auto r = by_size.equal_range(nx->second);
                for (auto it = r.first; it != r.second; ++it)
                    if (it->second == nx->first) { by_size.erase(it); break; }
                sz += nx->second;
                nx = by_off.erase(nx);
            }
            if (nx != by_off.begin()) {                           // merge with the previous block
                auto pv = std::prev(nx);
                if (pv->first + pv->second == off) {
                    auto r = by_size.equal_range(pv->second);
                    for (auto it = r.first; it != r.second; ++it)
                        if (it->second == pv->first) { by_size.erase(it); break; }
                    off = pv->first;
                    sz += pv->second;
                    by_off.erase(pv);
                }
            }
            if (off + sz == top) { top = off; return; }           // the block ends at the top
            by_off[off] = sz;
            by_size.emplace(sz, off);
        };
        auto acquire = [&](int64_t sz) -> int64_t {
            auto it = by_size.lower_bound(sz);
            if (it == by_size.end()) {
                const int64_t off = top;
                top += sz;
                P.ws_elems = std::max(P.ws_elems, top);
                return off;
            }
            const int64_t bsz = it->first, off = it->second;
            by_size.erase(it);
            by_off.erase(off);
            if (bsz > sz) {
                by_off[off + sz] = bsz - sz;
                by_size.emplace(bsz - sz, off + sz);
            }
            return off;
        };
        std::vector<int64_t> blk(nsuper, -1);
        for (int64_t L = 0; L < nlev; L++) {
            if (L > 0)
                for (int64_t s : ends[L - 1]) release(blk[s], need[s]);
            std::vector<int64_t> v = starts[L];
            std::stable_sort(v.begin(), v.end(), [&](int64_t x, int64_t y) { return need[x] > need[y]; });
            for (int64_t s : v) blk[s] = acquire(need[s]);
        }
        // the plan's own check: blocks live at the same level never overlap (sweep over the
        // levels keeping the live blocks ordered by offset; each insertion meets its neighbours)
        {
            std::map<int64_t, int64_t> live;   // offset -> end
            for (int64_t L = 0; L < nlev; L++) {
                if (L > 0)
                    for (int64_t s : ends[L - 1]) live.erase(blk[s]);
                for (int64_t s : starts[L]) {
                    const int64_t b = blk[s], e = b + need[s];
                    if (b < 0 || e > P.ws_elems) return fail(MF_BADARG, "internal: arena block out of range");
                    auto nx = live.lower_bound(b);
                    if (nx != live.end() && nx->first < e) return fail(MF_BADARG, "internal: arena blocks overlap");
                    if (nx != live.begin() && std::prev(nx)->second > b)
                        return fail(MF_BADARG, "internal: arena blocks overlap");
                    live[b] = e;
                }
            }
        }
        for (int64_t s = 0; s < nsuper; s++) {
            if (nf[s] > P.small_nf) {
                woff[s] = blk[s];
                cboff[s] = blk[s] + np[s] + np[s] * nf[s];
                cbld[s] = nf[s];
            } else {
                cboff[s] = blk[s] >= 0 ? blk[s] : 0;
                cbld[s] = std::max<int64_t>(nu[s], 1);
            }
        }
    }
    // original entries: V = As[qf,qf]; entry (iV, jV) goes to the front of column min(iV,jV)
    // (src/multifrontal.jl:216-231)
    std::vector<int64_t> qinv(n);
    for (int64_t j = 0; j < n; j++) qinv[P.qf[j]] = j;
    std::vector<int64_t> ocount(nsuper + 1, 0);
    struct OE { int32_t row, col, src; };
    std::vector<OE> oe(P.nnz);
    std::vector<int32_t> ofront(P.nnz);
    auto front_pos = [&](int64_t s, int64_t g) -> int64_t {
        if (g >= c1[s] && g < c1[s] + np[s]) return g - c1[s];
        const int32_t *b = upd.data() + uoff0[s], *e = b + nu[s];
        const int32_t *it = std::lower_bound(b, e, (int32_t)g);
        if (it == e || *it != g) return -1;
        return np[s] + (it - b);
    };
    for (int64_t jA = 0; jA < n; jA++) {
        for (int64_t p = Ap1[jA] - 1; p < Ap1[jA + 1] - 1; p++) {
            int64_t iA = Ai1[p] - 1;
            if (iA < 0 || iA >= n) return fail(MF_BADARG, "row index out of range");
            int64_t iV = qinv[iA], jV = qinv[jA];
            int64_t s, lr, lc;
            if (iV >= jV) {
                s = super_of[jV];
                lr = front_pos(s, iV);
                lc = jV - c1[s];
            } else {
                s = super_of[iV];
                lr = iV - c1[s];
                lc = front_pos(s, jV);
            }
            if (lr < 0 || lc < 0) return fail(MF_BADARG, "entry outside the symbolic structure");
            oe[p] = {(int32_t)lr, (int32_t)lc, (int32_t)p};
            ofront[p] = (int32_t)s;
            ocount[s + 1]++;
        }
    }
    for (int64_t s = 0; s < nsuper; s++) ocount[s + 1] += ocount[s];
    std::vector<OE> oes(P.nnz);
    {
        std::vector<int64_t> cur(ocount.begin(), ocount.end() - 1);
        for (int64_t p = 0; p < P.nnz; p++) oes[cur[ofront[p]]++] = oe[p];
    }
    std::vector<int32_t> orow(P.nnz), ocol(P.nnz), osrc(P.nnz), ocolptr(fcol[nsuper] + 1, 0);
    for (int64_t s = 0; s < nsuper; s++) {
        std::sort(oes.begin() + ocount[s], oes.begin() + ocount[s + 1], [](const OE &a, const OE &b) {
            return a.col != b.col ? a.col < b.col : a.row < b.row;
        });
        for (int64_t q = ocount[s]; q < ocount[s + 1]; q++) {
            orow[q] = oes[q].row;
            ocol[q] = oes[q].col;
            osrc[q] = oes[q].src;
            ocolptr[fcol[s] + oes[q].col + 1]++;
        }
    }
    // ocolptr: entries of front s are contiguous from ocount[s]; make per-column pointers global
    for (int64_t s = 0; s < nsuper; s++) {
        int64_t base = ocount[s];
        int64_t run = base;
        for (int64_t c = 0; c < nf[s]; c++) {
            int64_t cnt = ocolptr[fcol[s] + c + 1];
            ocolptr[fcol[s] + c] = (int32_t)run;
            run += cnt;
        }
    }
    ocolptr[fcol[nsuper]] = (int32_t)P.nnz;
    // child contributions per parent column, reverse child order within a column
    std::vector<int32_t> ecolptr(fcol[nsuper] + 1, 0);
    for (int64_t s = 0; s < nsuper; s++)
        for (int64_t c : kids[s])
            for (int64_t t = 0; t < nu[c]; t++) ecolptr[fcol[s] + rmap[uoff0[c] + t] + 1]++;
    for (int64_t q = 0; q < fcol[nsuper]; q++) ecolptr[q + 1] += ecolptr[q];
    std::vector<int2> ext(ecolptr[fcol[nsuper]]);
    std::vector<int32_t> eupd(ext.size());   // the same entries as offsets into the update vectors
    {
        std::vector<int32_t> cur(ecolptr.begin(), ecolptr.end() - 1);
        for (int64_t s = 0; s < nsuper; s++)
            for (auto it = kids[s].rbegin(); it != kids[s].rend(); ++it) {
                int64_t c = *it;
                for (int64_t t = 0; t < nu[c]; t++) {
                    int64_t col = fcol[s] + rmap[uoff0[c] + t];
                    eupd[cur[col]] = (int32_t)(uoff0[c] + t);
                    ext[cur[col]++] = make_int2((int)c, (int)t);
                }
            }
    }
    // the front position (index into ecolptr) each extend-add entry belongs to
    std::vector<int32_t> egpos(std::max<size_t>(eupd.size(), 1), 0);
    for (int64_t c = 0; c < fcol[nsuper]; c++)
        for (int32_t q = ecolptr[c]; q < ecolptr[c + 1]; q++) egpos[q] = (int32_t)c;
    // children in reverse order (the order src/multifrontal.jl:236 absorbs them)
    std::vector<int32_t> child_rev;
    std::vector<int64_t> child_off(nsuper);
    for (int64_t s = 0; s < nsuper; s++) {
        child_off[s] = (int64_t)child_rev.size();
        for (auto it = kids[s].rbegin(); it != kids[s].rend(); ++it) child_rev.push_back((int32_t)*it);
    }
    // per-supernode device records
    P.sn.resize(nsuper);
    for (int64_t s = 0; s < nsuper; s++) {
        SnDev &d = P.sn[s];
        d.foff = foff[s];
        d.u12off = u12off[s];
        d.woff = woff[s];
        d.cboff = cboff[s];
        d.cbld = (int32_t)cbld[s];
        d.c1 = (int32_t)c1[s];
        d.np = (int32_t)np[s];
        d.nu = (int32_t)nu[s];
        d.nf = (int32_t)nf[s];
        d.upd_off = (int32_t)uoff0[s];
        d.orig_off = (int32_t)ocount[s];
        d.norig = (int32_t)(ocount[s + 1] - ocount[s]);
        d.child_off = (int32_t)child_off[s];
        d.nchild = (int32_t)kids[s].size();
        d.fcol = (int32_t)fcol[s];
        d.pad = 0;
        d.ioff = ioff[s];
    }
    // flops
    for (int64_t s = 0; s < nsuper; s++) {
        double p = (double)np[s], u = (double)nu[s];
        double getrf = 0;
        for (int64_t k = 1; k <= np[s]; k++) getrf += (double)(np[s] - k) + 2.0 * (double)(np[s] - k) * (double)(np[s] - k);
        double g = 2.0 * p * u * u;
        P.flops += getrf + u * p * p + u * p * (p - 1) + g;
        P.gemm_flops += g;
        if (nf[s] >= 1024) P.gemm_flops_big += g;
    }
    // ---- launch schedule
    std::vector<std::vector<int64_t>> bylevel(nlev);
    for (int64_t s = 0; s < nsuper; s++) bylevel[level[s]].push_back(s);
    std::vector<int32_t> lists;
    std::vector<GemmProb> probs;
    std::vector<GemmTile> tiles;
    const double ebytes = (double)P.esize();
    std::vector<Launch> *gemm_dst = nullptr;
    // gemm launches: the problems of one call are grouped by the tile shape the policy picks
    // for their size bucket (one launch per shape used)
    const int32_t *policy = g_tile_policy[P.prec == MF_FP64 ? 0 : 1];
    std::function<void(std::vector<GemmProb> &, int, int)> add_gemm;
    add_gemm = [&](std::vector<GemmProb> &pv, int lvl, int schur) {
        if (pv.empty()) return;
        if (schur == 1) {
            // Schur complements of fronts with nf >= 1024 get launches of their own (variant 3),
            // so their MFMA utilisation is measured apart
            std::vector<GemmProb> big, rest;
            for (const GemmProb &g : pv) ((int64_t)g.m + g.k >= 1024 ? big : rest).push_back(g);
            if (!big.empty()) add_gemm(big, lvl, 3);
            if (rest.empty()) return;
            pv.swap(rest);
        }
        const int use = schur == 1 || schur == 3 ? 1 : 0;
        if (schur == 3 && P.prec == MF_BF16) {
            // BF16 Schur complements of the fronts with nf >= 1024: one launch packs the operand
            // tiles (scaled, rounded), the next multiplies them; pack tiles are reused launch by launch
            Launch Pk{}, L{};
            Pk.kind = L_PACK_BF16;
            L.kind = L_GEMM;
            L.variant = schur;
            L.tile = GEMM_TILE_BF16;
            Pk.prob_off = L.prob_off = (int32_t)probs.size();
            L.list_off = (int32_t)tiles.size();
            std::vector<int32_t> items;
            int64_t tcount = 0;
            for (size_t q = 0; q < pv.size(); q++) {
                GemmProb g = pv[q];
                const int64_t mt = (g.m + PK_ROWS - 1) / PK_ROWS, nt = (g.n + PK_ROWS - 1) / PK_ROWS,
                              kt = (g.k + PK_K - 1) / PK_K;
                g.apk = tcount;
                g.bpk = tcount + mt * kt;
                tcount += (mt + nt) * kt;
                for (int64_t r = 0; r < mt; r++)
                    for (int64_t c = 0; c < kt; c++) {
                        items.push_back((int32_t)q); items.push_back(0); items.push_back((int32_t)r); items.push_back((int32_t)c);
                    }
                for (int64_t r = 0; r < nt; r++)
                    for (int64_t c = 0; c < kt; c++) {
                        items.push_back((int32_t)q); items.push_back(1); items.push_back((int32_t)r); items.push_back((int32_t)c);
                    }
                for (int tn = 0; tn < g.n; tn += PK_ROWS)
                    for (int tm = 0; tm < g.m; tm += PK_ROWS) tiles.push_back({(int32_t)q, tm, tn, 0});
                probs.push_back(g);
                L.flops += 2.0 * g.m * (double)g.n * g.k;
                L.bytes += ((double)(mt + nt) * kt * PK_TILE * 2.0) + 2.0 * g.m * (double)g.n * 4.0;
                Pk.bytes += ((double)g.m * g.k + (double)g.k * g.n) * 4.0 + (double)(mt + nt) * kt * PK_TILE * 2.0;
            }
            P.pack_tiles = std::max(P.pack_tiles, tcount);
            while (lists.size() & 3) lists.push_back(0);
            Pk.list_off = (int32_t)lists.size();
            lists.insert(lists.end(), items.begin(), items.end());
            Pk.count = (int32_t)(items.size() / 4);
            Pk.level = L.level = lvl;
            L.count = (int32_t)(tiles.size() - L.list_off);
            {
                std::vector<int64_t> kw;
                for (const GemmProb &g : pv) kw.push_back(g.k);
                xcd_order_by_problem(tiles, L.list_off, kw);
            }
            gemm_dst->push_back(Pk);
            gemm_dst->push_back(L);
            return;
        }
        for (int tile = 0; tile < GEMM_NTILES; tile++) {
            std::vector<const GemmProb *> sel;
            for (const GemmProb &g : pv)
                if (policy[use * GEMM_NBUCKETS + gemm_bucket(g.m, g.n)] == tile) sel.push_back(&g);
            if (sel.empty()) continue;
            Launch L{};
            L.kind = L_GEMM;
            L.variant = schur;
            L.on_store = schur == 2 ? 1 : 0;   // the pivot-block inversions work in the store
            L.tile = tile;
            for (const GemmProb *g : sel) {
                // a triangular operand halves the products that can be nonzero (and its bytes)
                const bool ta = g->flags & (GEMM_A_LOWER | GEMM_A_UPPER), tb = g->flags & (GEMM_B_LOWER | GEMM_B_UPPER);
                L.flops += (ta || tb ? 1.0 : 2.0) * g->m * (double)g->n * g->k;
                L.bytes += ((double)g->m * g->k * (ta ? 0.5 : 1.0) + (double)g->k * g->n * (tb ? 0.5 : 1.0) +
                            2.0 * g->m * (double)g->n) * ebytes;
            }
            L.prob_off = (int32_t)probs.size();
            L.list_off = (int32_t)tiles.size();
            for (size_t q = 0; q < sel.size(); q++) {
                const GemmProb &g = *sel[q];
                for (int tn = 0; tn < g.n; tn += GEMM_TILE_N[tile])
                    for (int tm = 0; tm < g.m; tm += GEMM_TILE_M[tile]) tiles.push_back({(int32_t)q, tm, tn, 0});
                probs.push_back(g);
            }
            L.count = (int32_t)(tiles.size() - L.list_off);
            L.level = lvl;
            {
                std::vector<int64_t> kw;
                std::vector<char> bcyc;
                for (const GemmProb *g : sel) {
                    kw.push_back(g->k);
                    bcyc.push_back((g->flags & (GEMM_B_LOWER | GEMM_B_UPPER)) ? 1 : 0);
                }
                xcd_order_by_problem(tiles, L.list_off, kw, bcyc);
            }
            if (L.count > 0) gemm_dst->push_back(L);
        }
    };
    const int elem = (int)P.esize();
    // algorithmic bytes of assembling front s: the front written once, every child
    // contribution block read once, each original entry (value, scale, 3 int32 indices)
    auto asm_bytes = [&](int64_t s) {
        double b = (double)nf[s] * nf[s] * elem + 28.0 * (double)(ocount[s + 1] - ocount[s]);
        for (int64_t c : kids[s]) b += (double)nu[c] * nu[c] * elem + 4.0 * nu[c];
        return b;
    };
    auto front_flops = [&](int64_t s) {
        double p = (double)np[s], u = (double)nu[s], g = 0;
        for (int64_t k = 1; k <= np[s]; k++) g += (double)(np[s] - k) + 2.0 * (double)(np[s] - k) * (double)(np[s] - k);
        return g + u * p * p + u * p * (p - 1) + 2.0 * p * u * u;
    };
    for (int sp = 0; sp < 2; sp++) {
    std::vector<Launch> &sched = P.sched[sp];
    gemm_dst = &sched;
    for (int64_t lv = 0; lv < nlev; lv++) {
        std::vector<int64_t> small, large;
        for (int64_t s : bylevel[lv]) (nf[s] <= P.small_nf ? small : large).push_back(s);
        if (!large.empty()) P.has_large = true;
        if (!large.empty()) {
            // column assembly (int2 entries: keep 8-byte alignment)
            if (lists.size() & 1) lists.push_back(0);
            Launch L{};
            L.kind = L_ASM_COLS;
            L.list_off = (int32_t)lists.size();
            int64_t mnf = 0;
            for (int64_t s : large) {
                for (int64_t c = 0; c < nf[s]; c++) { lists.push_back((int32_t)s); lists.push_back((int32_t)c); }
                mnf = std::max(mnf, nf[s]);
                L.bytes += asm_bytes(s);
            }
            L.count = (int32_t)((lists.size() - L.list_off) / 2);
            L.shmem = (int32_t)(mnf * elem);
            L.level = (int32_t)lv;
            if (L.shmem > 160 * 1024) return fail(MF_BADARG, "front too large for column assembly");
            sched.push_back(L);
            Launch St{};   // appended after the level's factorization below
            St.kind = L_STORE;
            St.list_off = L.list_off;
            St.count = L.count;
            St.level = (int32_t)lv;
            for (int64_t s : large) St.bytes += 2.0 * ((double)nf[s] * np[s] + (double)np[s] * nu[s]) * elem;
            int64_t maxnp = 0;
            for (int64_t s : large) maxnp = std::max(maxnp, np[s]);
            if (sp == 1) {
                // Interchange-free blocks of SPEC_KB columns with the multiplier check, with a
                // look-ahead: after block k's trsm, one launch holds the next diagonal blocks
                // (each workgroup first applies block k's update to its own diagonal block,
                // then factors it), the rest of the next block's panel column and row, and the
                // remaining trailing update.  Every entry still receives the same products in
                // the same order.
                auto diag_fronts = [&](int64_t k) {
                    std::vector<int64_t> v;
                    for (int64_t s : large)
                        if (np[s] > k) v.push_back(s);
                    return v;
                };
                auto add_diag = [&](Launch &L, const std::vector<int64_t> &fr, int64_t k) {
                    for (size_t q = 0; q < fr.size(); q++) {
                        const int64_t s = fr[q];
                        lists.push_back((int32_t)s);
                        P.spec_slots = std::max<int64_t>(P.spec_slots, (int64_t)q + 1);
                        const double b = (double)std::min<int64_t>(SPEC_KB, np[s] - k);
                        L.flops += 2.0 * b * b * b / 3.0;
                        L.bytes += 2.0 * b * b * elem;
                    }
                };
                auto prob = [&](int64_t s, int64_t r0, int64_t c0, int64_t m, int64_t n, int64_t k, int64_t kb) {
                    GemmProb g{};
                    g.aoff = woff[s] + r0 + k * nf[s];
                    g.boff = woff[s] + k + c0 * nf[s];
                    g.coff = woff[s] + r0 + c0 * nf[s];
                    g.m = (int32_t)m;
                    g.n = (int32_t)n;
                    g.k = (int32_t)kb;
                    g.lda = g.ldb = g.ldc = (int32_t)nf[s];
                    return g;
                };
                {
                    Launch Dg{};
                    Dg.kind = L_SPEC_DIAG;
                    Dg.k = 0;
                    Dg.tau = P.pivot_tol;
                    Dg.level = (int32_t)lv;
                    Dg.list_off = (int32_t)lists.size();
                    add_diag(Dg, diag_fronts(0), 0);
                    Dg.count = (int32_t)(lists.size() - Dg.list_off);
                    sched.push_back(Dg);
                }
                for (int64_t k = 0; k < maxnp; k += SPEC_KB) {
                    const std::vector<int64_t> fr = diag_fronts(k);
                    std::vector<int32_t> items;
                    std::vector<GemmProb> g1v, g2v, dgv;
                    for (size_t q = 0; q < fr.size(); q++) {
                        const int64_t s = fr[q];
                        const int32_t slot = (int32_t)q;
                        const int64_t kb = std::min<int64_t>(SPEC_KB, np[s] - k), k2 = k + kb;
                        for (int64_t r0 = 0; r0 < nf[s] - k2; r0 += SPEC_ROWS) {
                            items.push_back((int32_t)s); items.push_back(0); items.push_back((int32_t)r0); items.push_back(slot);
                        }
                        for (int64_t c0 = 0; c0 < nf[s] - k2; c0 += SPEC_ROWS) {
                            items.push_back((int32_t)s); items.push_back(1); items.push_back((int32_t)c0); items.push_back(slot);
                        }
                        if (k2 >= np[s]) continue;
                        const int64_t k3 = k2 + std::min<int64_t>(SPEC_KB, np[s] - k2);
                        dgv.push_back(prob(s, k2, k2, k3 - k2, k3 - k2, k, kb));                            // next diagonal block
                        if (nf[s] > k3) g1v.push_back(prob(s, k3, k2, nf[s] - k3, k3 - k2, k, kb));         // panel column below it
                        if (nf[s] > k3) g1v.push_back(prob(s, k2, k3, k3 - k2, nf[s] - k3, k, kb));         // panel row
                        if (np[s] > k3) g2v.push_back(prob(s, k3, k3, nf[s] - k3, np[s] - k3, k, kb));      // rest of the pivot columns
                        if (np[s] > k3 && nu[s] > 0) g2v.push_back(prob(s, k3, np[s], np[s] - k3, nu[s], k, kb));  // rest of U12
                    }
                    if (!items.empty()) {
                        while (lists.size() & 3) lists.push_back(0);
                        Launch Tr{};
                        Tr.kind = L_SPEC_TRSM;
                        Tr.k = (int32_t)k;
                        Tr.tau = P.pivot_tol;
                        Tr.level = (int32_t)lv;
                        Tr.list_off = (int32_t)lists.size();
                        lists.insert(lists.end(), items.begin(), items.end());
                        Tr.count = (int32_t)(items.size() / 4);
                        for (size_t q = 0; q < items.size(); q += 4) {
                            int64_t s = items[q];
                            double b = (double)std::min<int64_t>(SPEC_KB, np[s] - k);
                            double cnt = (double)std::min<int64_t>(SPEC_ROWS, nf[s] - (k + (int64_t)b) - items[q + 2]);
                            Tr.flops += cnt * b * b;
                            Tr.bytes += 2.0 * cnt * b * elem;
                        }
                        sched.push_back(Tr);
                    }
                    const int64_t next = k + SPEC_KB;
                    const std::vector<int64_t> nfr = diag_fronts(next);
                    if (nfr.empty() && g2v.empty()) continue;
                    if (dgv.size() != nfr.size()) return fail(MF_BADARG, "internal: look-ahead blocks out of step");
                    g1v.insert(g1v.end(), g2v.begin(), g2v.end());
                    Launch C{};
                    C.kind = L_SPEC_DIAG_GEMM;
                    C.k = (int32_t)next;
                    C.tau = P.pivot_tol;
                    C.level = (int32_t)lv;
                    C.list_off = (int32_t)lists.size();
                    add_diag(C, nfr, next);
                    C.prob_off = (int32_t)probs.size();
                    C.tile_off = (int32_t)tiles.size();
                    for (size_t q = 0; q < g1v.size(); q++) {
                        const GemmProb &g = g1v[q];
                        for (int tn = 0; tn < g.n; tn += 64)
                            for (int tm = 0; tm < g.m; tm += 64) tiles.push_back({(int32_t)q, tm, tn, 0});
                        probs.push_back(g);
                        C.flops += 2.0 * g.m * (double)g.n * g.k;
                        C.bytes += ((double)g.m * g.k + (double)g.k * g.n + 2.0 * g.m * (double)g.n) * ebytes;
                    }
                    C.dprob = dgv.empty() ? -1 : (int32_t)g1v.size();
                    for (const GemmProb &g : dgv) {
                        probs.push_back(g);
                        C.flops += 2.0 * g.m * (double)g.n * g.k;
                        C.bytes += ((double)g.m * g.k + (double)g.k * g.n + 2.0 * g.m * (double)g.n) * ebytes;
                    }
                    xcd_order(tiles, C.tile_off);
                    C.ntiles = (int32_t)(tiles.size() - C.tile_off);
                    C.count = C.ntiles + (int32_t)nfr.size();
                    sched.push_back(C);
                }
            }
            for (int64_t k = 0; sp == 0 && k < maxnp; k += PANEL_KB) {
                Launch Pn{};
                Pn.kind = L_PANEL;
                Pn.k = (int32_t)k;
                Pn.list_off = (int32_t)lists.size();
                int64_t rows = 0;
                std::vector<GemmProb> pv;
                for (int64_t s : large) {
                    if (np[s] <= k) continue;
                    lists.push_back((int32_t)s);
                    rows = std::max(rows, np[s] - k);
                    int64_t kb = std::min<int64_t>(PANEL_KB, np[s] - k), k2 = k + kb;
                    {
                        double m = (double)(np[s] - k), b = (double)kb, right = (double)(nf[s] - k2);
                        for (int64_t j = 0; j < kb; j++)
                            Pn.flops += (m - j - 1) + 2.0 * (m - j - 1) * (b - j - 1);
                        Pn.flops += right * b * (b - 1) + (double)nu[s] * b * b;
                        Pn.bytes += 2.0 * m * b * elem;
                    }
                    if (k2 < np[s]) {
                        // rows [k2, nf) x cols [k2, np): pivot-column block incl. L21 rows
                        GemmProb g{};
                        g.aoff = woff[s] + k2 + k * nf[s];
                        g.boff = woff[s] + k + k2 * nf[s];
                        g.coff = woff[s] + k2 + k2 * nf[s];
                        g.m = (int32_t)(nf[s] - k2);
                        g.n = (int32_t)(np[s] - k2);
                        g.k = (int32_t)kb;
                        g.lda = g.ldb = g.ldc = (int32_t)nf[s];
                        pv.push_back(g);
                        if (nu[s] > 0) {  // rows [k2, np) x cols [np, nf): the U12 block
                            GemmProb h{};
                            h.aoff = woff[s] + k2 + k * nf[s];
                            h.boff = woff[s] + k + np[s] * nf[s];
                            h.coff = woff[s] + k2 + np[s] * nf[s];
                            h.m = (int32_t)(np[s] - k2);
                            h.n = (int32_t)nu[s];
                            h.k = (int32_t)kb;
                            h.lda = h.ldb = h.ldc = (int32_t)nf[s];
                            pv.push_back(h);
                        }
                    }
                }
                Pn.count = (int32_t)(lists.size() - Pn.list_off);
                Pn.variant = choose_panel_variant((int)rows);
                if (Pn.variant < 0) return fail(MF_BADARG, "pivot block too large for the panel kernel");
                Pn.level = (int32_t)lv;
                sched.push_back(Pn);
                // swaps, U rows and L21 columns of this panel, spread over many workgroups
                while (lists.size() & 3) lists.push_back(0);
                Launch Ap{};
                Ap.kind = L_PANEL_APPLY;
                Ap.k = (int32_t)k;
                Ap.list_off = (int32_t)lists.size();
                Ap.level = (int32_t)lv;
                for (int32_t q = Pn.list_off; q < Pn.list_off + Pn.count; q++) {
                    int64_t s = lists[q];
                    int64_t kb = std::min<int64_t>(PANEL_KB, np[s] - k);
                    for (int64_t c0 = 0; c0 < nf[s] - kb; c0 += 256) {
                        lists.push_back((int32_t)s); lists.push_back(0); lists.push_back((int32_t)c0); lists.push_back(0);
                    }
                    for (int64_t r0 = 0; r0 < nu[s]; r0 += 256) {
                        lists.push_back((int32_t)s); lists.push_back(1); lists.push_back((int32_t)r0); lists.push_back(0);
                    }
                    Ap.bytes += 2.0 * ((double)kb * (nf[s] - kb) + (double)nu[s] * kb) * elem;
                }
                Ap.count = (int32_t)((lists.size() - Ap.list_off) / 4);
                if (Ap.count > 0) sched.push_back(Ap);
                add_gemm(pv, (int)lv, 0);
            }
            std::vector<GemmProb> sv;
            for (int64_t s : large) {
                if (nu[s] == 0) continue;
                GemmProb g{};
                g.aoff = woff[s] + np[s];
                g.boff = woff[s] + np[s] * nf[s];
                g.coff = woff[s] + np[s] + np[s] * nf[s];
                g.m = g.n = (int32_t)nu[s];
                g.k = (int32_t)np[s];
                g.lda = g.ldb = g.ldc = (int32_t)nf[s];
                sv.push_back(g);
            }
            add_gemm(sv, (int)lv, 1);
            sched.push_back(St);
        }
        if (!small.empty()) {
            Launch L{};
            L.kind = L_FUSED_SMALL;
            L.list_off = (int32_t)lists.size();
            int64_t mnf = 0, mnp = 0;
            for (int64_t s : small) {
                lists.push_back((int32_t)s);
                mnf = std::max(mnf, nf[s]);
                mnp = std::max(mnp, np[s]);
                L.flops += front_flops(s);
                L.bytes += asm_bytes(s);
            }
            L.count = (int32_t)small.size();
            L.shmem = (int32_t)(mnf * mnf * elem + mnp * 4);
            L.level = (int32_t)lv;
            if (L.shmem > 160 * 1024) return fail(MF_BADARG, "small_nf too large for LDS");
            sched.push_back(L);
        }
    }
    // recursive inversion of the pivot-block factors of the large fronts (after every level).
    // The fronts of the top level are inverted in launches of their own (P.inv_top, built once)
    // that the first solve runs beside the sweeps of the levels below, which do not need them.
    for (int part = 0; part < (sp == 0 ? 2 : 1); part++) {
        const bool top = part == 1;
        std::vector<Launch> &idst = top ? P.inv_top : sched;
        gemm_dst = &idst;
        std::vector<int32_t> items;
        for (int64_t s = 0; s < nsuper; s++)
            if (ioff[s] >= 0 && (level[s] == nlev - 1) == top)
                for (int64_t b0 = 0; b0 < np[s]; b0 += INV_LEAF) {
                    items.push_back((int32_t)s); items.push_back((int32_t)b0); items.push_back(0); items.push_back(0);
                }
        if (!items.empty()) {
            while (lists.size() & 3) lists.push_back(0);
            Launch Lf{};
            Lf.kind = L_INV_LEAF;
            Lf.list_off = (int32_t)lists.size();
            lists.insert(lists.end(), items.begin(), items.end());
            Lf.count = (int32_t)(items.size() / 4);
            Lf.level = (int32_t)nlev;
            Lf.bytes = 0;
            for (size_t q = 0; q < items.size(); q += 4) {
                double b = (double)std::min<int64_t>(INV_LEAF, np[items[q]] - items[q + 1]);
                Lf.flops += 2.0 * b * b * b / 3.0;
                Lf.bytes += 3.0 * b * b * elem;
            }
            idst.push_back(Lf);
            for (int64_t half = INV_LEAF;; half *= 2) {
                std::vector<GemmProb> t1, t2;
                int64_t toff = pe;   // temps live after the inverses
                for (int64_t s = 0; s < nsuper; s++) {
                    if (ioff[s] < 0 || (level[s] == nlev - 1) != top) continue;
                    const int64_t f = foff[s], Li = ioff[s], Ui = ioff[s] + np[s] * np[s], ld = np[s];
                    for (int64_t a0 = 0; a0 + half < np[s]; a0 += 2 * half) {
                        int64_t m = a0 + half, b = std::min(a0 + 2 * half, np[s]);
                        int64_t s1 = m - a0, s2 = b - m;
                        // L: T = -M inv(L1) (s2 x s1), X21 = inv(L2) T
                        GemmProb g{};
                        g.aoff = f + m + a0 * nf[s]; g.lda = (int32_t)nf[s];
                        g.boff = Li + a0 + a0 * ld; g.ldb = (int32_t)ld;
                        g.coff = toff; g.ldc = (int32_t)s2;
                        g.m = (int32_t)s2; g.n = (int32_t)s1; g.k = (int32_t)s1; g.flags = GEMM_SET | GEMM_B_LOWER;
                        t1.push_back(g);
                        GemmProb h{};
                        h.aoff = Li + m + m * ld; h.lda = (int32_t)ld;
                        h.boff = toff; h.ldb = (int32_t)s2;
                        h.coff = Li + m + a0 * ld; h.ldc = (int32_t)ld;
                        h.m = (int32_t)s2; h.n = (int32_t)s1; h.k = (int32_t)s2; h.flags = GEMM_SET | GEMM_ADD | GEMM_A_LOWER;
                        t2.push_back(h);
                        toff += s2 * s1;
                        // U: T = -N inv(U2) (s1 x s2), X12 = inv(U1) T
                        GemmProb u{};
                        u.aoff = f + a0 + m * nf[s]; u.lda = (int32_t)nf[s];
                        u.boff = Ui + m + m * ld; u.ldb = (int32_t)ld;
                        u.coff = toff; u.ldc = (int32_t)s1;
                        u.m = (int32_t)s1; u.n = (int32_t)s2; u.k = (int32_t)s2; u.flags = GEMM_SET | GEMM_B_UPPER;
                        t1.push_back(u);
                        GemmProb v{};
                        v.aoff = Ui + a0 + a0 * ld; v.lda = (int32_t)ld;
                        v.boff = toff; v.ldb = (int32_t)s1;
                        v.coff = Ui + a0 + m * ld; v.ldc = (int32_t)ld;
                        v.m = (int32_t)s1; v.n = (int32_t)s2; v.k = (int32_t)s1; v.flags = GEMM_SET | GEMM_ADD | GEMM_A_UPPER;
                        t2.push_back(v);
                        toff += s1 * s2;
                    }
                }
                if (t1.empty()) break;
                P.tmp_elems = std::max(P.tmp_elems, toff - pe);
                add_gemm(t1, (int)nlev, 2);
                add_gemm(t2, (int)nlev, 2);
            }
        }
        gemm_dst = &sched;
    }
    }
    P.store_elems = pe + P.tmp_elems;
    // solve lists: all fronts per level
    std::vector<int32_t> slist;
    P.lvl_off.assign(nlev, 0);
    P.lvl_cnt.assign(nlev, 0);
    auto emit4 = [&](int32_t a0, int32_t a1, int32_t a2, int32_t a3) {
        slist.push_back(a0); slist.push_back(a1); slist.push_back(a2); slist.push_back(a3);
    };
    for (int64_t lv = 0; lv < nlev; lv++) {
        // The fused one-workgroup-per-front sweeps pay off when a level has many fronts (times
        // the matrices of a batch).  A level with only a few fronts beyond one wave sends all
        // its fronts through the split pipeline, whose launches spread each front over many
        // workgroups; large fronts (stored inverses or nf > FRONT_MAXNF) always take it.
        int64_t nmid = 0;
        for (int64_t s : bylevel[lv])
            if (ioff[s] < 0 && nf[s] > 64 && nf[s] <= FRONT_MAXNF_HOST) nmid++;
        const bool split = nmid > 0 && nmid * std::max(P.nbatch, 1) < SOLVE_FUSE_MIN;
        std::vector<int64_t> big;
        auto is_big = [&](int64_t s) { return ioff[s] >= 0 || nf[s] > FRONT_MAXNF_HOST; };
        P.lvl_s64o.push_back((int32_t)slist.size());
        for (int64_t s : bylevel[lv])
            if (!split && !is_big(s) && nf[s] <= 64) slist.push_back((int32_t)s);
        P.lvl_s64c.push_back((int32_t)slist.size() - P.lvl_s64o.back());
        P.lvl_s1ko.push_back((int32_t)slist.size());
        int64_t mnf1k = 0;
        for (int64_t s : bylevel[lv])
            if (!split && !is_big(s) && nf[s] > 64) {
                slist.push_back((int32_t)s);
                mnf1k = std::max(mnf1k, nf[s]);
            }
        P.lvl_s1k_small.push_back(mnf1k <= 256 ? 1 : 0);
        P.lvl_s1kc.push_back((int32_t)slist.size() - P.lvl_s1ko.back());
        for (int64_t s : bylevel[lv])
            if (split || is_big(s)) big.push_back(s);
        P.lvl_off[lv] = (int32_t)slist.size();
        for (int64_t s : big) slist.push_back((int32_t)s);
        P.lvl_cnt[lv] = (int32_t)big.size();
        // forward: rows t < nu of L21 against columns a < np; backward: rows i < np of U12
        // against columns t < nu.  Partial slots are reused level by level.
        while (slist.size() & 3) slist.push_back(0);
        P.lvl_go.push_back((int32_t)slist.size());
        for (int64_t s : big)
            for (int64_t p0 = 0; p0 < nf[s]; p0 += 256) emit4((int32_t)s, (int32_t)p0, 0, 0);
        P.lvl_gc.push_back((int32_t)((slist.size() - P.lvl_go.back()) / 4));
        // dir 0: L21 (rows nu, cols np); 1: U12 (rows np, cols nu); 2: inv(L11) and 3: inv(U11)
        // of the large fronts (rows np, cols np; chunks entirely outside the triangle skipped).
        // U12 of a root (nu = 0) has no chunks; its sum still sets x = y.
        int32_t ntri = 0;
        int64_t maxnp = 0;
        for (int64_t s : big) {
            ntri += ioff[s] < 0 ? 1 : 0;
            maxnp = std::max(maxnp, np[s]);
        }
        P.lvl_ntri.push_back(ntri);
        // L21 of a level whose pivot blocks fit four column chunks: one launch, no split-K
        const bool fdirect = maxnp <= 4 * SOLVE_CK_HOST;
        while (slist.size() & 3) slist.push_back(0);
        P.lvl_fdo.push_back((int32_t)slist.size());
        if (fdirect)
            for (int64_t s : big)
                for (int64_t r0 = 0; r0 < nu[s]; r0 += 64) emit4((int32_t)s, (int32_t)r0, 0, 0);
        P.lvl_fdc.push_back((int32_t)((slist.size() - P.lvl_fdo.back()) / 4));
        for (int dir = 0; dir < 4; dir++) {
            int64_t slot = 0;
            std::vector<int32_t> sums;
            while (slist.size() & 3) slist.push_back(0);
            int32_t po = (int32_t)slist.size();
            for (int64_t s : big) {
                if (dir >= 2 && ioff[s] < 0) continue;
                if (dir == 0 && fdirect) continue;
                int64_t nrows = dir == 0 ? nu[s] : np[s], ncols = dir == 1 ? nu[s] : np[s];
                for (int64_t r0 = 0; r0 < nrows; r0 += 256) {
                    int64_t nch = 0;
                    for (int64_t c0 = 0; c0 < ncols; c0 += SOLVE_CK_HOST) {
                        if (dir == 2 && c0 > r0 + 255) continue;            // above the lower triangle
                        if (dir == 3 && c0 + SOLVE_CK_HOST <= r0) continue;  // below the upper triangle
                        emit4((int32_t)s, (int32_t)r0, (int32_t)c0, (int32_t)(slot + nch));
                        nch++;
                    }
                    const int64_t dst = (dir == 0 ? uoff0[s] : c1[s]) + r0;
                    const int64_t rows = std::min<int64_t>(256, nrows - r0);
                    const int64_t to_zb = dir == 1 && ioff[s] >= 0 ? 1 : 0;
                    sums.push_back((int32_t)dst); sums.push_back((int32_t)(rows | to_zb << 16));
                    sums.push_back((int32_t)nch); sums.push_back((int32_t)slot);
                    slot += nch;
                }
            }
            int32_t pc = (int32_t)((slist.size() - po) / 4);
            int32_t so = (int32_t)slist.size();
            slist.insert(slist.end(), sums.begin(), sums.end());
            int32_t sc = (int32_t)(sums.size() / 4);
            switch (dir) {
            case 0: P.lvl_fpo.push_back(po); P.lvl_fpc.push_back(pc); P.lvl_fso.push_back(so); P.lvl_fsc.push_back(sc); break;
            case 1: P.lvl_bpo.push_back(po); P.lvl_bpc.push_back(pc); P.lvl_bso.push_back(so); P.lvl_bsc.push_back(sc); break;
            case 2: P.lvl_ilpo.push_back(po); P.lvl_ilpc.push_back(pc); P.lvl_ilso.push_back(so); P.lvl_ilsc.push_back(sc); break;
            default: P.lvl_iupo.push_back(po); P.lvl_iupc.push_back(pc); P.lvl_iuso.push_back(so); P.lvl_iusc.push_back(sc); break;
            }
            P.part_slots = std::max(P.part_slots, slot);
        }
        double lf = 0, uf = 0, lb = 0, ub = 0, vec = 0;
        for (int64_t s : bylevel[lv]) {
            double p = (double)np[s], f = (double)nf[s];
            lf += 2.0 * (p * (p - 1) / 2 + (f - p) * p);   // unit-lower L11 + L21
            uf += 2.0 * (p * (p + 1) / 2 + p * (f - p));   // U11 + U12
            lb += (p * (p - 1) / 2 + (f - p) * p) * elem;
            ub += (p * (p + 1) / 2 + p * (f - p)) * elem;
            vec += f;
        }
        P.lvl_lflops.push_back(lf);
        P.lvl_uflops.push_back(uf);
        P.lvl_lbytes.push_back(lb);
        P.lvl_ubytes.push_back(ub);
        P.lvl_vec.push_back(vec);
        P.lvl_allo.push_back((int32_t)slist.size());
        std::vector<int64_t> nfs;
        for (int64_t s : bylevel[lv]) {
            slist.push_back((int32_t)s);
            P.maxnp = std::max(P.maxnp, np[s]);
            nfs.push_back(nf[s]);
        }
        P.lvl_allc.push_back((int32_t)bylevel[lv].size());
        {
            int32_t ch = 8;
            if (!nfs.empty()) {
                std::sort(nfs.begin(), nfs.end());
                const int64_t p75 = nfs[std::min(nfs.size() - 1, (nfs.size() * 3) / 4)];
                while (ch < 64 && ch < p75) ch *= 2;
            }
            P.lvl_csch.push_back(ch);
        }
    }
    // CSC export offsets (src/multifrontal.jl:73-104, :277-317)
    P.Lp.assign(n + 1, 0);
    P.Up.assign(n + 1, 0);
    {
        std::vector<int64_t> ucount(n, 0);
        for (int64_t s = 0; s < nsuper; s++) {
            for (int64_t a = 0; a < np[s]; a++) {
                int64_t j = c1[s] + a;
                P.Lp[j + 1] = 1 + (np[s] - 1 - a) + nu[s];
                ucount[j] += a + 1;
            }
            for (int64_t t = 0; t < nu[s]; t++) ucount[upd[uoff0[s] + t]] += np[s];
        }
        for (int64_t j = 0; j < n; j++) P.Lp[j + 1] += P.Lp[j];
        for (int64_t j = 0; j < n; j++) P.Up[j + 1] = P.Up[j] + ucount[j];
        P.uoff.assign(P.nupd, 0);
        std::vector<int64_t> cur(P.Up.begin(), P.Up.end() - 1);
        for (int64_t s = 0; s < nsuper; s++)
            for (int64_t t = 0; t < nu[s]; t++) {
                int64_t J = upd[uoff0[s] + t];
                P.uoff[uoff0[s] + t] = cur[J];
                cur[J] += np[s];
            }
        P.nnzL = P.Lp[n];
        P.nnzU = P.Up[n];
    }
    // CSR of A for the residual: rows ascending within, column order kept per row
    std::vector<int32_t> rp(n + 1, 0), ci(P.nnz), vmap(P.nnz), arow(P.nnz);
    for (int64_t p = 0; p < P.nnz; p++) { arow[p] = (int32_t)(Ai1[p] - 1); rp[arow[p] + 1]++; }
    for (int64_t i = 0; i < n; i++) rp[i + 1] += rp[i];
    {
        std::vector<int32_t> cur(rp.begin(), rp.end() - 1);
        for (int64_t j = 0; j < n; j++)
            for (int64_t p = Ap1[j] - 1; p < Ap1[j + 1] - 1; p++) {
                int32_t q = cur[arow[p]]++;
                ci[q] = (int32_t)j;
                vmap[q] = (int32_t)p;
            }
    }
    std::vector<int32_t> qf32(n);
    for (int64_t j = 0; j < n; j++) qf32[j] = (int32_t)P.qf[j];

    if (!device) return MF_OK;   // mf_plan_estimate: sizes and schedule only
    // ---- device uploads
    HIPCHK(upload(P.d_sn, P.sn));
    HIPCHK(upload(P.d_upd, upd));
    HIPCHK(upload(P.d_rmap, rmap));
    HIPCHK(upload(P.d_child, child_rev));
    HIPCHK(upload(P.d_orow, orow));
    HIPCHK(upload(P.d_ocol, ocol));
    HIPCHK(upload(P.d_osrc, osrc));
    HIPCHK(upload(P.d_ocolptr, ocolptr));
    HIPCHK(upload(P.d_ecolptr, ecolptr));
    {
        std::vector<ExtRec> er(std::max<size_t>(ext.size(), 1));
        for (size_t q = 0; q < ext.size(); q++) {
            const int64_t c = ext[q].x, t = ext[q].y;
            er[q] = {cboff[c] + t * cbld[c], (int32_t)uoff0[c], (int32_t)nu[c]};
        }
        HIPCHK(upload(P.d_ext, er));
    }
    HIPCHK(upload(P.d_eupd, eupd));
    HIPCHK(upload(P.d_egpos, egpos));
    HIPCHK(upload(P.d_arow, arow));
    HIPCHK(upload(P.d_lists, lists));
    HIPCHK(upload(P.d_probs, probs));
    HIPCHK(upload(P.d_tiles, tiles));
    HIPCHK(upload(P.d_solvelist, slist));
    HIPCHK(upload(P.d_qf, qf32));
    {
        std::vector<int32_t> qinv(n);
        for (int64_t j = 0; j < n; j++) qinv[P.qf[j]] = (int32_t)j;
        HIPCHK(upload(P.d_qinv, qinv));
    }
    HIPCHK(upload(P.d_csr_rp, rp));
    HIPCHK(upload(P.d_csr_ci, ci));
    HIPCHK(upload(P.d_csr_vmap, vmap));
    HIPCHK(upload(P.d_uoff, P.uoff));
    // the column-parallel sweeps read up to CS_CH_HOST elements past a front's factor blocks
    HIPCHK(P.d_store.alloc(((size_t)P.store_elems * P.nbatch + CS_CH_HOST) * P.esize()));
    HIPCHK(hipMemset(P.d_store.p, 0, ((size_t)P.store_elems * P.nbatch + CS_CH_HOST) * P.esize()));
    HIPCHK(P.d_ws.alloc((size_t)std::max<int64_t>(P.ws_elems, 16) * P.esize() * P.nbatch));
    HIPCHK(hipMemset(P.d_ws.p, 0, (size_t)std::max<int64_t>(P.ws_elems, 16) * P.esize() * P.nbatch));
    HIPCHK(P.d_redo.alloc((size_t)4 * P.nbatch));
    HIPCHK(P.d_pack.alloc((size_t)std::max<int64_t>(P.pack_tiles, 1) * PK_TILE * 2 * P.nbatch));
    HIPCHK(P.d_pexp.alloc((size_t)std::max<int64_t>(P.pack_tiles, 1) * 4 * P.nbatch));
    HIPCHK(P.d_prow.alloc((size_t)n * 4 * P.nbatch));
    HIPCHK(P.d_info.alloc((size_t)4 * P.nbatch));
    HIPCHK(P.d_pinfo.alloc((size_t)4 * PINFO_STRIDE * nsuper * P.nbatch));
    HIPCHK(P.d_flag.alloc((size_t)4 * P.nbatch));
    HIPCHK(P.d_spec.alloc((size_t)std::max<int64_t>(1, P.spec_slots) * SPEC_SLOT * P.esize() * P.nbatch));
    HIPCHK(P.d_ax.alloc((size_t)P.nnz * 8 * P.nbatch));
    HIPCHK(P.d_rs.alloc((size_t)n * 8 * P.nbatch));
    HIPCHK(P.d_anorm.alloc((size_t)8 * P.nbatch));
    HIPCHK(P.d_npart.alloc((size_t)8 * 256 * P.nbatch));
    return MF_OK;
}

static int plan_options(mf_plan &P, const mf_options *opt)
{
    if (opt) {
        if (opt->precision < MF_FP64 || opt->precision > MF_BF16) return fail(MF_BADARG, "unknown precision");
        P.prec = opt->precision;
        if (opt->small_nf > 0) P.small_nf = opt->small_nf;
        P.use_graph = opt->use_graph;
        P.nbatch = opt->nbatch > 0 ? opt->nbatch : 1;
        if (opt->pivoting != MF_PIVOT_AUTO && opt->pivoting != MF_PIVOT_PARTIAL)
            return fail(MF_BADARG, "unknown pivoting mode");
        P.pivoting = opt->pivoting;
        if (opt->pivot_tol < 0.f || opt->pivot_tol > 1.f) return fail(MF_BADARG, "pivot_tol must be in [0, 1]");
        P.pivot_tol = opt->pivot_tol;
        if (opt->solve_path < MF_SOLVE_AUTO || opt->solve_path > MF_SOLVE_COLUMNS)
            return fail(MF_BADARG, "unknown solve_path");
        P.solve_path = opt->solve_path;
    }
    if (P.pivot_tol == 0.f) P.pivot_tol = P.prec == MF_FP64 ? 1.f : 0.1f;
    if (P.small_nf > 128) return fail(MF_BADARG, "small_nf must be <= 128");
    return MF_OK;
}

static int plan_args_ok(int64_t n, const int64_t *qf, int64_t nsuper, const int64_t *sstart, const int64_t *upd_ptr,
                        const int64_t *child_ptr, const int64_t *Ap)
{
    if (n <= 0) return fail(MF_DIM, "matrix must have n >= 1");
    if (!qf || !sstart || !upd_ptr || !child_ptr || !Ap || nsuper <= 0 || nsuper > n)
        return fail(MF_BADARG, "missing plan arguments");
    return MF_OK;
}

extern "C" int mf_plan_create(int64_t n, const int64_t *qf, int64_t nsuper, const int64_t *sstart,
                              const int64_t *upd_ptr, const int64_t *upd_idx, const int64_t *child_ptr,
                              const int64_t *child_idx, const int64_t *Ap, const int64_t *Ai, const mf_options *opt,
                              mf_plan **out)
{
    if (!out) return fail(MF_BADARG, "out is NULL");
    *out = nullptr;
    int rc = plan_args_ok(n, qf, nsuper, sstart, upd_ptr, child_ptr, Ap);
    if (rc != MF_OK) return rc;
    std::unique_ptr<mf_plan> P(new (std::nothrow) mf_plan);
    if (!P) return fail(MF_OOM, "host allocation failed");
    if ((rc = plan_options(*P, opt)) != MF_OK) return rc;
    HIPCHK(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
    try {
        rc = build_plan(*P, n, qf, nsuper, sstart, upd_ptr, upd_idx, child_ptr, child_idx, Ap, Ai, true);
    } catch (const std::bad_alloc &) {
        return fail(MF_OOM, "host allocation failed while building the plan");
    }
    if (rc != MF_OK) return rc;
    if (!P->inv_top.empty()) {
        HIPCHK(hipStreamCreateWithFlags(&P->side, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&P->ev_fork, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&P->ev_inv, hipEventDisableTiming));
    }
    *out = P.release();
    return MF_OK;
}

static void fill_stats(const mf_plan *P, mf_plan_stats *st)
{
    st->factor_flops = P->flops;
    st->gemm_flops = P->gemm_flops;
    st->gemm_flops_big = P->gemm_flops_big;
    st->front_bytes = ((double)(P->store_elems + P->ws_elems) * (double)P->esize() + (double)P->pack_tiles * PK_TILE * 2.0) *
                      P->nbatch;
    st->store_bytes = (double)P->store_elems * (double)P->esize() * P->nbatch;
    st->arena_bytes = (double)P->ws_elems * (double)P->esize() * P->nbatch;
    st->resident_front_bytes = (double)P->ws_sum_elems * (double)P->esize() * P->nbatch;
    st->last_redone = P->last_redone;
    st->nsuper = P->nsuper;
    st->nlevels = P->nlevels;
    st->launches = (int64_t)P->sched[P->last_sched].size();
    st->last_schedule = P->last_sched;
    st->max_nf = P->maxnf;
}

extern "C" int mf_plan_estimate(int64_t n, const int64_t *qf, int64_t nsuper, const int64_t *sstart,
                                const int64_t *upd_ptr, const int64_t *upd_idx, const int64_t *child_ptr,
                                const int64_t *child_idx, const int64_t *Ap, const int64_t *Ai, const mf_options *opt,
                                mf_plan_stats *st)
{
    if (!st) return fail(MF_BADARG, "NULL stats");
    int rc = plan_args_ok(n, qf, nsuper, sstart, upd_ptr, child_ptr, Ap);
    if (rc != MF_OK) return rc;
    std::unique_ptr<mf_plan> P(new (std::nothrow) mf_plan);
    if (!P) return fail(MF_OOM, "host allocation failed");
    if ((rc = plan_options(*P, opt)) != MF_OK) return rc;
    try {
        rc = build_plan(*P, n, qf, nsuper, sstart, upd_ptr, upd_idx, child_ptr, child_idx, Ap, Ai, false);
    } catch (const std::bad_alloc &) {
        return fail(MF_OOM, "host allocation failed while building the plan");
    }
    if (rc != MF_OK) return rc;
    P->last_sched = P->has_large && P->pivoting == MF_PIVOT_AUTO ? 1 : 0;
    fill_stats(P.get(), st);
    st->last_redone = 0;
    return MF_OK;
}

extern "C" void mf_plan_destroy(mf_plan *P) { delete P; }

extern "C" int mf_nnz(const mf_plan *P, int64_t *nnzL, int64_t *nnzU)
{
    if (!P) return fail(MF_BADARG, "NULL plan");
    if (nnzL) *nnzL = P->nnzL;
    if (nnzU) *nnzU = P->nnzU;
    return MF_OK;
}

extern "C" int mf_plan_get_stats(const mf_plan *P, mf_plan_stats *st)
{
    if (!P || !st) return fail(MF_BADARG, "NULL argument");
    fill_stats(P, st);
    return MF_OK;
}

// ------------------------------------------------------------------ factor
// Graph slots: the interchange-free schedule, the pivoted schedule for every matrix, and the
// pivoted schedule for only the matrices whose interchange-free attempt failed.
enum { G_SPEC = 0, G_PIV_ALL = 1, G_PIV_REDO = 2 };

static FactorArgs factor_args(mf_plan &P, const double *d_ax, const double *d_rs, bool redo);

// The top level's pivot-block inversions on stream st, for every factored matrix.
template <typename T>
static hipError_t launch_inv_top(mf_plan &P, int nmat, hipStream_t st)
{
    const FactorArgs a = factor_args(P, nullptr, nullptr, false);
    for (const Launch &L : P.inv_top) {
        hipError_t e = launch_factor_step<T>(L, a, P.d_lists.as<int32_t>(), P.d_probs.as<GemmProb>(),
                                             P.d_tiles.as<GemmTile>(), nmat, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <typename T>
static int run_factor(mf_plan &P, const double *d_ax, const double *d_rs, int nmat, int sp, bool redo)
{
    FactorArgs a = factor_args(P, d_ax, d_rs, redo);
    HIPCHK(launch_fill_i32_masked(P.d_info.as<int32_t>(), INT32_MAX, a.only, nmat, P.stream));
    if (!redo) HIPCHK(launch_fill_i32(P.d_flag.as<int32_t>(), 0, nmat, P.stream));
    for (const Launch &L : P.sched[sp])
        HIPCHK(profiled(P, launch_class(L), L.flops * nmat, L.bytes * nmat, [&] {
            return launch_factor_step<T>(L, a, P.d_lists.as<int32_t>(), P.d_probs.as<GemmProb>(),
                                         P.d_tiles.as<GemmTile>(), nmat, P.stream);
        }));
    if (P.prof) {
        // timed launch by launch: the top inversions stay in the factorization
        for (const Launch &L : P.inv_top)
            HIPCHK(profiled(P, launch_class(L), L.flops * nmat, L.bytes * nmat, [&] {
                return launch_factor_step<T>(L, a, P.d_lists.as<int32_t>(), P.d_probs.as<GemmProb>(),
                                             P.d_tiles.as<GemmTile>(), nmat, P.stream);
            }));
        P.inv_pending = false;
    } else {
        P.inv_pending = !P.inv_top.empty();
    }
    if (!redo)
        HIPCHK(launch_anorm(P.d_csr_rp.as<int32_t>(), P.d_csr_vmap.as<int32_t>(), d_ax, P.nnz, P.n,
                            P.d_npart.as<double>(), P.d_anorm.as<double>(), nmat, P.stream));
    return MF_OK;
}

static FactorArgs factor_args(mf_plan &P, const double *d_ax, const double *d_rs, bool redo)
{
    FactorArgs a{};
    a.sn = P.d_sn.as<SnDev>();
    a.orig_row = P.d_orow.as<int32_t>();
    a.orig_col = P.d_ocol.as<int32_t>();
    a.orig_src = P.d_osrc.as<int32_t>();
    a.orig_colptr = P.d_ocolptr.as<int32_t>();
    a.child_list = P.d_child.as<int32_t>();
    a.rmap = P.d_rmap.as<int32_t>();
    a.ext_colptr = P.d_ecolptr.as<int32_t>();
    a.ext = P.d_ext.as<ExtRec>();
    a.arow = P.d_arow.as<int32_t>();
    a.ax = d_ax;
    a.rs = d_rs;
    a.ws = P.d_ws.p;
    a.store = P.d_store.p;
    a.prow = P.d_prow.as<int32_t>();
    a.info = P.d_info.as<int32_t>();
    a.pinfo = P.d_pinfo.as<int32_t>();
    a.flag = P.d_flag.as<int32_t>();
    // a redo reads the failed attempts' flags (a copy: the pivoted kernels leave them alone,
    // but the mask must not change under a captured graph's feet either way)
    a.only = redo ? P.d_redo.as<int32_t>() : nullptr;
    a.spec = P.d_spec.p;
    a.spec_stride = P.spec_slots * SPEC_SLOT;
    a.pack = P.d_pack.p;
    a.pexp = P.d_pexp.as<int32_t>();
    a.pack_tiles = P.pack_tiles;
    a.pinfo_stride = (int64_t)PINFO_STRIDE * P.nsuper;
    a.ws_stride = P.ws_elems;
    a.store_stride = P.store_elems;
    a.ax_stride = P.nnz;
    a.rs_stride = P.n;
    a.n_stride = P.n;
    return a;
}

// Enqueues schedule sp for every matrix, or (redo) the pivoted schedule for the matrices whose
// interchange-free attempt failed, directly or through a captured graph.
static int enqueue_schedule(mf_plan *P, const double *d_ax, const double *d_rs, int nmat, int sp, bool redo)
{
    const int rsg = d_rs != nullptr;
    const int gi = sp == 1 ? G_SPEC : redo ? G_PIV_REDO : G_PIV_ALL;
    auto run = [&]() {
        return P->prec == MF_FP64 ? run_factor<double>(*P, d_ax, d_rs, nmat, sp, redo)
                                  : run_factor<float>(*P, d_ax, d_rs, nmat, sp, redo);
    };
    if (P->use_graph && !P->prof) {
        // inputs are always read from the plan's own buffers, so a captured graph stays valid
        if (!P->graph[gi] || P->graph_nmat[gi] != nmat || rsg != P->graph_rs[gi]) {
            if (P->graph[gi]) { (void)hipGraphExecDestroy(P->graph[gi]); P->graph[gi] = nullptr; }
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(P->stream, hipStreamCaptureModeThreadLocal));
            const int rc = run();
            const hipError_t ec = hipStreamEndCapture(P->stream, &g);
            if (rc != MF_OK || ec != hipSuccess) {
                if (ec == hipSuccess && g) (void)hipGraphDestroy(g);
                if (rc != MF_OK) return rc;
                HIPCHK(ec);
            }
            const hipError_t ei = hipGraphInstantiate(&P->graph[gi], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) P->graph[gi] = nullptr;
            HIPCHK(ei);
            P->graph_nmat[gi] = nmat;
            P->graph_rs[gi] = rsg;
        }
        HIPCHK(hipGraphLaunch(P->graph[gi], P->stream));
        return MF_OK;
    }
    return run();
}

static int factor_common(mf_plan *P, const double *d_ax, const double *d_rs, int nmat, int32_t check,
                         int64_t *info_col)
{
    int rc;
    if (P->side_busy) {   // a solve that ended early may have left the top inversions running
        HIPCHK(hipStreamSynchronize(P->side));
        P->side_busy = false;
    }
    std::vector<int32_t> info(nmat), flag(nmat, 0);
    // Every call first tries the interchange-free schedule (when it applies); the matrices whose
    // multiplier check failed are then factored again on the pivoted schedule, from the assembly
    // on, while the others keep their factors.  The policy is per call and per matrix: one
    // matrix that needs interchanges changes nothing for the next call or the rest of a batch.
    const bool spec = P->has_large && P->pivoting == MF_PIVOT_AUTO;
    int nbad = nmat;
    if (spec) {
        if ((rc = enqueue_schedule(P, d_ax, d_rs, nmat, 1, false)) != MF_OK) return rc;
        HIPCHK(hipMemcpyAsync(flag.data(), P->d_flag.p, (size_t)nmat * 4, hipMemcpyDeviceToHost, P->stream));
        HIPCHK(hipStreamSynchronize(P->stream));
        nbad = 0;
        for (int m = 0; m < nmat; m++) nbad += flag[m] != 0;
        if (nbad > 0 && nbad < nmat) {
            HIPCHK(hipMemcpyAsync(P->d_redo.p, P->d_flag.p, (size_t)nmat * 4, hipMemcpyDeviceToDevice, P->stream));
            if ((rc = enqueue_schedule(P, d_ax, d_rs, nmat, 0, true)) != MF_OK) return rc;
        }
    }
    if (nbad == nmat)
        if ((rc = enqueue_schedule(P, d_ax, d_rs, nmat, 0, false)) != MF_OK) return rc;
    P->last_sched = spec && nbad == 0 ? 1 : 0;
    P->last_redone = spec ? nbad : nmat;
    P->rs_given = d_rs != nullptr;
    HIPCHK(hipMemcpyAsync(info.data(), P->d_info.p, (size_t)nmat * 4, hipMemcpyDeviceToHost, P->stream));
    HIPCHK(hipStreamSynchronize(P->stream));
    if (P->prof) HIPCHK(P->ev_collect());
    P->factored = true;
    P->nmat_factored = nmat;
    bool sing = false;
    for (int m = 0; m < nmat; m++) {
        int64_t c = info[m] == INT32_MAX ? 0 : (int64_t)info[m] + 1;
        if (info_col) info_col[m] = c;
        if (c) sing = true;
    }
    if (check && sing) {
        for (int m = 0; m < nmat; m++)
            if (info[m] != INT32_MAX) {
                g_err = "SingularException(" + std::to_string((long long)info[m] + 1) + ")";
                break;
            }
        return MF_SINGULAR;
    }
    return MF_OK;
}

extern "C" int mf_factor_device(mf_plan *P, const double *d_Ax, const double *d_Rs, int32_t check, int64_t *info_col)
{
    if (!P || !d_Ax) return fail(MF_BADARG, "NULL argument");
    // the plan keeps its own copy: the refinement residual reads A and Rs again later
    HIPCHK(hipMemcpyAsync(P->d_ax.p, d_Ax, (size_t)P->nnz * 8, hipMemcpyDeviceToDevice, P->stream));
    if (d_Rs) HIPCHK(hipMemcpyAsync(P->d_rs.p, d_Rs, (size_t)P->n * 8, hipMemcpyDeviceToDevice, P->stream));
    return factor_common(P, P->d_ax.as<double>(), d_Rs ? P->d_rs.as<double>() : nullptr, 1, check, info_col);
}

extern "C" int mf_factor(mf_plan *P, const double *Ax, const double *Rs, int32_t check, int64_t *info_col)
{
    if (!P || !Ax) return fail(MF_BADARG, "NULL argument");
    HIPCHK(hipMemcpyAsync(P->d_ax.p, Ax, (size_t)P->nnz * 8, hipMemcpyHostToDevice, P->stream));
    if (Rs) HIPCHK(hipMemcpyAsync(P->d_rs.p, Rs, (size_t)P->n * 8, hipMemcpyHostToDevice, P->stream));
    return factor_common(P, P->d_ax.as<double>(), Rs ? P->d_rs.as<double>() : nullptr, 1, check, info_col);
}

static int refactor_batch(mf_plan *P, int64_t nmat, const double *Ax_batch, const double *Rs_batch,
                          int64_t *info_col, hipMemcpyKind kind)
{
    if (!P || !Ax_batch) return fail(MF_BADARG, "NULL argument");
    if (nmat < 1 || nmat > P->nbatch) return fail(MF_BADARG, "nmat exceeds the plan's nbatch");
    HIPCHK(hipMemcpyAsync(P->d_ax.p, Ax_batch, (size_t)P->nnz * 8 * nmat, kind, P->stream));
    if (Rs_batch) HIPCHK(hipMemcpyAsync(P->d_rs.p, Rs_batch, (size_t)P->n * 8 * nmat, kind, P->stream));
    return factor_common(P, P->d_ax.as<double>(), Rs_batch ? P->d_rs.as<double>() : nullptr, (int)nmat, 0, info_col);
}
extern "C" int mf_refactor_batch(mf_plan *P, int64_t nmat, const double *Ax_batch, const double *Rs_batch,
                                 int64_t *info_col)
{
    return refactor_batch(P, nmat, Ax_batch, Rs_batch, info_col, hipMemcpyHostToDevice);
}
extern "C" int mf_refactor_batch_device(mf_plan *P, int64_t nmat, const double *d_Ax_batch, const double *d_Rs_batch,
                                        int64_t *info_col)
{
    return refactor_batch(P, nmat, d_Ax_batch, d_Rs_batch, info_col, hipMemcpyDeviceToDevice);
}

// ------------------------------------------------------------------ solve
static bool use_columns(const mf_plan &P, int64_t nrhs, int nmat)
{
    if (P.solve_path == MF_SOLVE_COLUMNS) return true;
    if (P.solve_path == MF_SOLVE_FRONTS) return false;
    return nrhs * nmat >= CS_MIN_COLS && P.maxnp <= CS_MAXNP;
}

// Column-parallel solve of every column of the batch: two launches per level.
template <typename T>
static int solve_columns(mf_plan &P, const double *d_rhs, int64_t nrhs, int accumulate, int nmat)
{
    const int64_t n = P.n, C = nrhs * nmat;
    const double *rs = P.rs_given ? P.d_rs.as<double>() : nullptr;
    HIPCHK(launch_cs_prep<T>(d_rhs, n, n * nrhs, rs, n, P.d_qinv.as<int32_t>(), reinterpret_cast<T *>(P.sb.w0), n,
                             (int)nrhs, C, P.stream));
    SolveArgs a{};
    a.sn = P.d_sn.as<SnDev>();
    a.upd = P.d_upd.as<int32_t>();
    a.store = P.d_store.p;
    a.prow = P.d_prow.as<int32_t>();
    a.w0 = P.sb.w0;
    a.y = P.sb.y;
    a.x = P.sb.x;
    a.uv = P.sb.uv;
    a.zbuf = P.sb.zbuf;
    a.ecolptr = P.d_ecolptr.as<int32_t>();
    a.eupd = P.d_eupd.as<int32_t>();
    a.egpos = P.d_egpos.as<int32_t>();
    a.n = n;
    a.nupd = P.nupd;
    a.nrhs = nrhs;
    a.store_stride = P.store_elems;
    a.n_stride = n;
    const int32_t *sl = P.d_solvelist.as<int32_t>();
    const double e = (double)P.esize();
    for (int64_t lv = 0; lv < P.nlevels; lv++)
        HIPCHK(profiled(P, 5, P.lvl_lflops[lv] * C, (P.lvl_lbytes[lv] * nmat + 3.0 * e * P.lvl_vec[lv] * C), [&] {
            return launch_cs_level<T>(1, a, sl, P.lvl_allo[lv], P.lvl_allc[lv], C, (int)nrhs, P.lvl_csch[lv],
                                      P.stream);
        }));
    for (int64_t lv = P.nlevels - 1; lv >= 0; lv--)
        HIPCHK(profiled(P, 6, P.lvl_uflops[lv] * C, (P.lvl_ubytes[lv] * nmat + 3.0 * e * P.lvl_vec[lv] * C), [&] {
            return launch_cs_level<T>(0, a, sl, P.lvl_allo[lv], P.lvl_allc[lv], C, (int)nrhs, P.lvl_csch[lv],
                                      P.stream);
        }));
    HIPCHK(launch_cs_finish<T>(reinterpret_cast<const T *>(P.sb.x), P.d_qinv.as<int32_t>(), P.sb.X, n, n * nrhs, n,
                               (int)nrhs, C, accumulate, P.stream));
    return MF_OK;
}

template <typename T>
static int solve_once(mf_plan &P, const double *d_rhs, int64_t nrhs, int accumulate, int nmat, bool inv)
{
    if (use_columns(P, nrhs, nmat)) return solve_columns<T>(P, d_rhs, nrhs, accumulate, nmat);
    if (inv) {
        // the top level's inversions beside the sweeps below it, joined before its level
        HIPCHK(hipEventRecord(P.ev_fork, P.stream));
        HIPCHK(hipStreamWaitEvent(P.side, P.ev_fork, 0));
        HIPCHK(launch_inv_top<T>(P, P.nmat_factored, P.side));
        HIPCHK(hipEventRecord(P.ev_inv, P.side));
    }
    // d_rhs: n x nrhs per matrix (ld n); result into P.sb.X (ld n), overwritten or accumulated
    const int64_t n = P.n;
    const double *rs = P.rs_given ? P.d_rs.as<double>() : nullptr;
    HIPCHK(launch_prep_rhs<T>(d_rhs, n, rs, n, P.d_qf.as<int32_t>(), reinterpret_cast<T *>(P.sb.w0), n, nrhs, n * nrhs, nmat,
                              P.stream));
    SolveArgs a{};
    a.sn = P.d_sn.as<SnDev>();
    a.child_list = P.d_child.as<int32_t>();
    a.rmap = P.d_rmap.as<int32_t>();
    a.upd = P.d_upd.as<int32_t>();
    a.store = P.d_store.p;
    a.prow = P.d_prow.as<int32_t>();
    a.w0 = P.sb.w0;
    a.y = P.sb.y;
    a.x = P.sb.x;
    a.uv = P.sb.uv;
    a.part = P.sb.part;
    a.zbuf = P.sb.zbuf;
    a.ecolptr = P.d_ecolptr.as<int32_t>();
    a.eupd = P.d_eupd.as<int32_t>();
    a.n = n;
    a.nupd = P.nupd;
    a.nrhs = nrhs;
    a.store_stride = P.store_elems;
    a.n_stride = n;
    a.part_stride = std::max<int64_t>(P.part_slots, 1) * 256 * nrhs;
    const double e = (double)P.esize();
    const int32_t *sl = P.d_solvelist.as<int32_t>();
    // one launch for a level's one-wave and workgroup fronts when the one-wave items are few
    // (many of them are better off without the workgroup fronts' LDS footprint)
    const int64_t nzc = (nrhs + 7) / 8;
    auto merge_ok = [&](int64_t wcount) { return wcount * nmat * nzc <= 4096; };
    for (int64_t lv = 0; lv < P.nlevels; lv++) {
        const double vb = 3.0 * e * P.lvl_vec[lv] * nrhs;
        if (inv && lv == P.nlevels - 1) HIPCHK(hipStreamWaitEvent(P.stream, P.ev_inv, 0));
        if (P.lvl_s64c[lv] > 0 && P.lvl_s1kc[lv] > 0 && merge_ok(P.lvl_s64c[lv])) {
            HIPCHK(profiled(P, 5, 0.0, 0.0, [&] {
                return launch_solve_merged<T>(1, P.lvl_s1k_small[lv], a, sl, P.lvl_s1ko[lv], P.lvl_s1kc[lv],
                                              P.lvl_s64o[lv], P.lvl_s64c[lv], nmat, P.stream);
            }));
        } else {
            if (P.lvl_s64c[lv] > 0)
                HIPCHK(profiled(P, 5, 0.0, 0.0, [&] {
                    return launch_solve_level<T>(11, a, sl, P.lvl_s64o[lv], P.lvl_s64c[lv], nmat, P.stream);
                }));
            if (P.lvl_s1kc[lv] > 0)
                HIPCHK(profiled(P, 5, 0.0, 0.0, [&] {
                    return launch_solve_level<T>(P.lvl_s1k_small[lv] ? 15 : 12, a, sl, P.lvl_s1ko[lv],
                                                 P.lvl_s1kc[lv], nmat, P.stream);
                }));
        }
        if (P.lvl_cnt[lv] == 0) continue;
        HIPCHK(profiled(P, 5, 0.0, vb * nmat, [&] {
            return launch_solve_level<T>(10, a, sl, P.lvl_go[lv], P.lvl_gc[lv], nmat, P.stream);
        }));
        if (P.lvl_ntri[lv] > 0)
            HIPCHK(profiled(P, 5, 0.0, 0.0, [&] {
                return launch_solve_level<T>(0, a, sl, P.lvl_off[lv], P.lvl_cnt[lv], nmat, P.stream);
            }));
        if (P.lvl_ilsc[lv] > 0) {
            HIPCHK(profiled(P, 5, 0.0, 0.0, [&] {
                return launch_solve_level<T>(6, a, sl, P.lvl_ilpo[lv], P.lvl_ilpc[lv], nmat, P.stream);
            }));
            HIPCHK(profiled(P, 5, 0.0, 0.0, [&] {
                return launch_solve_level<T>(7, a, sl, P.lvl_ilso[lv], P.lvl_ilsc[lv], nmat, P.stream);
            }));
        }
        if (P.lvl_fdc[lv] > 0)
            HIPCHK(profiled(P, 5, P.lvl_lflops[lv] * nrhs * nmat, P.lvl_lbytes[lv] * nmat, [&] {
                return launch_solve_level<T>(17, a, sl, P.lvl_fdo[lv], P.lvl_fdc[lv], nmat, P.stream);
            }));
        if (P.lvl_fpc[lv] > 0) {
            HIPCHK(profiled(P, 5, P.lvl_lflops[lv] * nrhs * nmat, P.lvl_lbytes[lv] * nmat, [&] {
                return launch_solve_level<T>(1, a, sl, P.lvl_fpo[lv], P.lvl_fpc[lv], nmat, P.stream);
            }));
            HIPCHK(profiled(P, 5, 0.0, 0.0, [&] {
                return launch_solve_level<T>(2, a, sl, P.lvl_fso[lv], P.lvl_fsc[lv], nmat, P.stream);
            }));
        }
    }
    for (int64_t lv = P.nlevels - 1; lv >= 0; lv--) {
        const double vb = 3.0 * e * P.lvl_vec[lv] * nrhs;
        if (P.lvl_s64c[lv] > 0 && P.lvl_s1kc[lv] > 0 && merge_ok(P.lvl_s64c[lv])) {
            HIPCHK(profiled(P, 6, 0.0, 0.0, [&] {
                return launch_solve_merged<T>(0, P.lvl_s1k_small[lv], a, sl, P.lvl_s1ko[lv], P.lvl_s1kc[lv],
                                              P.lvl_s64o[lv], P.lvl_s64c[lv], nmat, P.stream);
            }));
        } else {
            if (P.lvl_s64c[lv] > 0)
                HIPCHK(profiled(P, 6, 0.0, 0.0, [&] {
                    return launch_solve_level<T>(13, a, sl, P.lvl_s64o[lv], P.lvl_s64c[lv], nmat, P.stream);
                }));
            if (P.lvl_s1kc[lv] > 0)
                HIPCHK(profiled(P, 6, 0.0, 0.0, [&] {
                    return launch_solve_level<T>(P.lvl_s1k_small[lv] ? 16 : 14, a, sl, P.lvl_s1ko[lv],
                                                 P.lvl_s1kc[lv], nmat, P.stream);
                }));
        }
        if (P.lvl_cnt[lv] == 0) continue;
        if (P.lvl_bpc[lv] > 0)
            HIPCHK(profiled(P, 6, P.lvl_uflops[lv] * nrhs * nmat, (P.lvl_ubytes[lv] + vb) * nmat, [&] {
                return launch_solve_level<T>(3, a, sl, P.lvl_bpo[lv], P.lvl_bpc[lv], nmat, P.stream);
            }));
        HIPCHK(profiled(P, 6, 0.0, 0.0, [&] {
            return launch_solve_level<T>(4, a, sl, P.lvl_bso[lv], P.lvl_bsc[lv], nmat, P.stream);
        }));
        if (P.lvl_iusc[lv] > 0) {
            HIPCHK(profiled(P, 6, 0.0, 0.0, [&] {
                return launch_solve_level<T>(8, a, sl, P.lvl_iupo[lv], P.lvl_iupc[lv], nmat, P.stream);
            }));
            HIPCHK(profiled(P, 6, 0.0, 0.0, [&] {
                return launch_solve_level<T>(9, a, sl, P.lvl_iuso[lv], P.lvl_iusc[lv], nmat, P.stream);
            }));
        }
        if (P.lvl_ntri[lv] > 0)
            HIPCHK(profiled(P, 6, 0.0, 0.0, [&] {
                return launch_solve_level<T>(5, a, sl, P.lvl_off[lv], P.lvl_cnt[lv], nmat, P.stream);
            }));
    }
    HIPCHK(launch_finish_x<T>(reinterpret_cast<const T *>(P.sb.x), P.d_qf.as<int32_t>(), P.sb.X, n, n, nrhs, accumulate,
                              n * nrhs, nmat, P.stream));
    return MF_OK;
}

static int solve_core(mf_plan *P, int64_t nrhs, int nmat, int32_t refine_steps, double bwd_tol,
                      mf_solve_report *rep)
{
    // B already in P->sb.B (n x nrhs x nmat); result left in P->sb.X
    const int64_t n = P->n;
    auto t0 = std::chrono::steady_clock::now();
    int rc;
    // the first solve after a factorization computes the top level's inverses beside its sweeps
    // (timed runs and column sweeps, which read no inverse: on the stream first, or not at all)
    bool inv = P->inv_pending && !use_columns(*P, nrhs, nmat);
    if (inv && P->prof) {
        HIPCHK(P->prec == MF_FP64 ? launch_inv_top<double>(*P, P->nmat_factored, P->stream)
                                  : launch_inv_top<float>(*P, P->nmat_factored, P->stream));
        P->inv_pending = inv = false;
    }
    auto eager = [&](const double *rhs, int acc, bool with_inv) {
        return P->prec == MF_FP64 ? solve_once<double>(*P, rhs, nrhs, acc, nmat, with_inv)
                                  : solve_once<float>(*P, rhs, nrhs, acc, nmat, with_inv);
    };
    // With use_graph, each sweep pair (the solve from B, the correction from R, the solve from B
    // with the top inversions beside it) is captured once per (nrhs, nmat, buffers) and
    // replayed: the same kernels in the same order, so the same bits.
    auto once = [&](const double *rhs, int acc) -> int {
        const bool with_inv = inv && acc == 0;
        int r0;
        if (!P->use_graph || P->prof || P->sb_scratch) {
            r0 = eager(rhs, acc, with_inv);
            if (with_inv) {
                P->side_busy = true;   // even on an error part of it may have been enqueued
                if (r0 == MF_OK) P->inv_pending = inv = false;
            }
            return r0;
        }
        mf_plan::SolveGraph &G = P->sgraph[with_inv ? 2 : acc];
        if (!G.exec || G.nrhs != nrhs || G.nmat != nmat || G.rhs != rhs || G.w0 != P->sb.w0) {
            if (G.exec) { (void)hipGraphExecDestroy(G.exec); G.exec = nullptr; }
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(P->stream, hipStreamCaptureModeThreadLocal));
            const int r = eager(rhs, acc, with_inv);
            const hipError_t ec = hipStreamEndCapture(P->stream, &g);
            if (r != MF_OK || ec != hipSuccess) {
                if (ec == hipSuccess && g) (void)hipGraphDestroy(g);
                if (r != MF_OK) return r;
                HIPCHK(ec);
            }
            const hipError_t ei = hipGraphInstantiate(&G.exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) G.exec = nullptr;
            HIPCHK(ei);
            G.nrhs = nrhs;
            G.nmat = nmat;
            G.rhs = rhs;
            G.w0 = P->sb.w0;
        }
        HIPCHK(hipGraphLaunch(G.exec, P->stream));
        if (with_inv) { P->inv_pending = inv = false; P->side_busy = true; }
        return MF_OK;
    };
    auto norms = [&](std::vector<double> &h) -> int {
        HIPCHK(launch_residual(P->d_csr_rp.as<int32_t>(), P->d_csr_ci.as<int32_t>(), P->d_csr_vmap.as<int32_t>(),
                               P->d_ax.as<double>(), P->nnz, P->sb.X, n, P->sb.B, n,
                               P->sb.R, n, nrhs, n * nrhs, n * nrhs, nmat, P->stream));
        HIPCHK(launch_norms(P->sb.R, P->sb.X, P->sb.B, n, nrhs, P->sb.npart, P->sb.norms, nmat, P->stream));
        h.resize((size_t)5 * nrhs * nmat + nmat);
        HIPCHK(hipMemcpyAsync(h.data(), P->sb.norms, (size_t)5 * nrhs * nmat * 8, hipMemcpyDeviceToHost, P->stream));
        HIPCHK(hipMemcpyAsync(h.data() + 5 * nrhs * nmat, P->d_anorm.p, (size_t)nmat * 8, hipMemcpyDeviceToHost,
                              P->stream));
        HIPCHK(hipStreamSynchronize(P->stream));
        return MF_OK;
    };
    auto bwd = [&](const std::vector<double> &h, double &be, double &r2) {
        be = 0;
        r2 = 0;
        for (int m = 0; m < nmat; m++) {
            double an = h[5 * nrhs * nmat + m];
            for (int64_t c = 0; c < nrhs; c++) {
                const double *o = &h[(m * nrhs + c) * 5];
                double den = an * o[1] + o[2];
                be = std::max(be, den > 0 ? o[0] / den : o[0]);
                r2 = std::max(r2, o[4] > 0 ? std::sqrt(o[3] / o[4]) : std::sqrt(o[3]));
            }
        }
    };
    if ((rc = once(P->sb.B, 0)) != MF_OK) return rc;
    int it = 0, conv = 0;
    std::vector<double> h;
    double be = 0, r2 = 0;
    bool fresh = false;   // h holds the norms of the current solution
    for (; it < refine_steps; it++) {
        if (bwd_tol > 0) {
            if ((rc = norms(h)) != MF_OK) return rc;
            bwd(h, be, r2);
            fresh = true;
            if (be <= bwd_tol) { conv = 1; break; }
        } else {
            HIPCHK(launch_residual(P->d_csr_rp.as<int32_t>(), P->d_csr_ci.as<int32_t>(),
                                   P->d_csr_vmap.as<int32_t>(), P->d_ax.as<double>(), P->nnz, P->sb.X, n,
                                   P->sb.B, n, P->sb.R, n, nrhs, n * nrhs, n * nrhs, nmat,
                                   P->stream));
        }
        if ((rc = once(P->sb.R, 1)) != MF_OK) return rc;
        fresh = false;
    }
    if (rep) {
        if (!fresh) {
            if ((rc = norms(h)) != MF_OK) return rc;
            bwd(h, be, r2);
        }
        if (bwd_tol > 0 && be <= bwd_tol) conv = 1;
        rep->iterations = it;
        rep->converged = conv;
        rep->backward_error = be;
        rep->residual_2 = r2;
    } else {
        HIPCHK(hipStreamSynchronize(P->stream));
    }
    P->side_busy = false;   // joined into the stream before the top level's sweep, now synchronized
    auto t1 = std::chrono::steady_clock::now();
    if (rep) rep->seconds = std::chrono::duration<double>(t1 - t0).count();
    if (P->prof) {
        HIPCHK(hipStreamSynchronize(P->stream));
        HIPCHK(P->ev_collect());
    }
    return MF_OK;
}

// Carves the solve buffers for (nrhs, nmat) from `base` (256-byte aligned pieces) into *b;
// returns the bytes needed.  base == nullptr: size only.
static size_t solve_layout(const mf_plan &P, int64_t nrhs, int nmat, unsigned char *base, SolveBufs *b)
{
    size_t off = 0;
    auto take = [&](size_t bytes) -> void * {
        void *p = base ? base + off : nullptr;
        off += (bytes + 255) / 256 * 256;
        return p;
    };
    const size_t e = P.esize(), nv = (size_t)P.n * nrhs * nmat;
    SolveBufs t;
    t.w0 = take(nv * e);
    t.y = take(nv * e);
    t.x = take(nv * e);
    t.uv = take((size_t)std::max<int64_t>(P.nupd, 1) * nrhs * nmat * e);
    t.part = take((size_t)std::max<int64_t>(P.part_slots, 1) * 256 * nrhs * nmat * e);
    t.zbuf = take(nv * e);
    t.B = static_cast<double *>(take(nv * 8));
    t.X = static_cast<double *>(take(nv * 8));
    t.R = static_cast<double *>(take(nv * 8));
    t.norms = static_cast<double *>(take((size_t)5 * nrhs * nmat * 8));
    t.npart = static_cast<double *>(take((size_t)5 * 64 * nrhs * nmat * 8));
    if (b) *b = t;
    return off;
}

static int ensure_solve_buffers(mf_plan *P, int64_t nrhs, int nmat)
{
    HIPCHK(P->d_solve.alloc(solve_layout(*P, nrhs, nmat, nullptr, nullptr)));
    solve_layout(*P, nrhs, nmat, P->d_solve.as<unsigned char>(), &P->sb);
    return MF_OK;
}

extern "C" int mf_solve_device(mf_plan *P, int64_t nrhs, const double *d_B, int64_t ldb, double *d_X, int64_t ldx,
                               int32_t refine_steps, double bwd_tol, mf_solve_report *rep)
{
    if (!P || !d_B || !d_X) return fail(MF_BADARG, "NULL argument");
    if (!P->factored) return fail(MF_BADARG, "plan has no factorization");
    if (nrhs < 1 || ldb < P->n || ldx < P->n || refine_steps < 0) return fail(MF_DIM, "bad RHS dimensions");
    int rc = ensure_solve_buffers(P, nrhs, 1);
    if (rc != MF_OK) return rc;
    const size_t n8 = (size_t)P->n * 8;
    HIPCHK(hipMemcpy2DAsync(P->sb.B, n8, d_B, (size_t)ldb * 8, n8, nrhs, hipMemcpyDeviceToDevice, P->stream));
    if ((rc = solve_core(P, nrhs, 1, refine_steps, bwd_tol, rep)) != MF_OK) return rc;
    HIPCHK(hipMemcpy2DAsync(d_X, (size_t)ldx * 8, P->sb.X, n8, n8, nrhs, hipMemcpyDeviceToDevice, P->stream));
    HIPCHK(hipStreamSynchronize(P->stream));
    return MF_OK;
}

extern "C" int mf_solve(mf_plan *P, int64_t nrhs, const double *B, int64_t ldb, double *X, int64_t ldx,
                        int32_t refine_steps, double bwd_tol, mf_solve_report *rep)
{
    if (!P || !B || !X) return fail(MF_BADARG, "NULL argument");
    if (!P->factored) return fail(MF_BADARG, "plan has no factorization");
    if (nrhs < 1 || ldb < P->n || ldx < P->n || refine_steps < 0) return fail(MF_DIM, "bad RHS dimensions");
    int rc = ensure_solve_buffers(P, nrhs, 1);
    if (rc != MF_OK) return rc;
    const size_t n8 = (size_t)P->n * 8;
    HIPCHK(hipMemcpy2DAsync(P->sb.B, n8, B, (size_t)ldb * 8, n8, nrhs, hipMemcpyHostToDevice, P->stream));
    if ((rc = solve_core(P, nrhs, 1, refine_steps, bwd_tol, rep)) != MF_OK) return rc;
    HIPCHK(hipMemcpy2DAsync(X, (size_t)ldx * 8, P->sb.X, n8, n8, nrhs, hipMemcpyDeviceToHost, P->stream));
    HIPCHK(hipStreamSynchronize(P->stream));
    return MF_OK;
}

// ------------------------------------------------------------------ allocation-free solve
// The scratch of mf_solve_device_noalloc: a file-scope device array of MF_NOALLOC_SCRATCH
// bytes, one copy on each GPU the library runs on, held until the process ends.  In host code
// `scratch` names a host-side stand-in, not a device address: the address of the current
// device's copy comes from hipGetSymbolAddress, once per device, and only that pointer is
// handed to kernels and copies.  Every call on one GPU shares the array and leaves its
// contents behind; nothing in the solve reads a scratch byte before writing it, so nothing is
// cleared.  Calls that overlap on one GPU (two plans on two host threads) race on it.
static constexpr size_t MF_NOALLOC_SCRATCH = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[MF_NOALLOC_SCRATCH];
static constexpr int MF_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch[MF_MAX_DEVICES] = {};

static int scratch_address(void **out)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MF_MAX_DEVICES) return fail(MF_BADARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(MF_HIP, "hipGetSymbolAddress gave no address for the solve scratch");
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return MF_OK;
}

extern "C" int mf_solve_device_noalloc(mf_plan *P, int64_t nrhs, const double *d_B, int64_t ldb, double *d_X,
                                       int64_t ldx, int32_t refine_steps, double bwd_tol, mf_solve_report *rep)
{
    if (!P || !d_B || !d_X) return fail(MF_BADARG, "NULL argument");
    if (!P->factored) return fail(MF_BADARG, "plan has no factorization");
    if (nrhs < 1 || ldb < P->n || ldx < P->n || refine_steps < 0) return fail(MF_DIM, "bad RHS dimensions");
    if (P->prof) return fail(MF_BADARG, "profiling creates events: turn it off for the allocation-free solve");
    const size_t need = solve_layout(*P, nrhs, 1, nullptr, nullptr);
    if (need > MF_NOALLOC_SCRATCH)
        return fail(MF_BADARG, "this solve needs " + std::to_string(need) + " bytes of scratch; the allocation-free "
                    "entry point has " + std::to_string(MF_NOALLOC_SCRATCH) + " (use mf_solve_device)");
    void *base = nullptr;
    int rc = scratch_address(&base);
    if (rc != MF_OK) return rc;
    const SolveBufs own = P->sb;
    solve_layout(*P, nrhs, 1, static_cast<unsigned char *>(base), &P->sb);
    P->sb_scratch = true;   // no graph capture: instantiating one allocates
    const size_t n8 = (size_t)P->n * 8;
    hipError_t e = hipMemcpy2DAsync(P->sb.B, n8, d_B, (size_t)ldb * 8, n8, nrhs, hipMemcpyDeviceToDevice, P->stream);
    if (e == hipSuccess) {
        rc = solve_core(P, nrhs, 1, refine_steps, bwd_tol, rep);
        if (rc == MF_OK) {
            e = hipMemcpy2DAsync(d_X, (size_t)ldx * 8, P->sb.X, n8, n8, nrhs, hipMemcpyDeviceToDevice, P->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(P->stream);
        }
    }
    P->sb = own;
    P->sb_scratch = false;
    if (rc != MF_OK) return rc;
    HIPCHK(e);
    return MF_OK;
}

static int solve_batch(mf_plan *P, int64_t nmat, int64_t nrhs, const double *B, double *X, int32_t refine_steps,
                       double bwd_tol, mf_solve_report *rep, hipMemcpyKind in_kind, hipMemcpyKind out_kind)
{
    if (!P || !B || !X) return fail(MF_BADARG, "NULL argument");
    if (!P->factored || nmat != P->nmat_factored) return fail(MF_BADARG, "solve_batch needs a matching refactor_batch");
    if (nrhs < 1 || refine_steps < 0) return fail(MF_DIM, "bad RHS dimensions");
    int rc = ensure_solve_buffers(P, nrhs, (int)nmat);
    if (rc != MF_OK) return rc;
    size_t bytes = (size_t)P->n * nrhs * nmat * 8;
    HIPCHK(hipMemcpyAsync(P->sb.B, B, bytes, in_kind, P->stream));
    if ((rc = solve_core(P, nrhs, (int)nmat, refine_steps, bwd_tol, rep)) != MF_OK) return rc;
    HIPCHK(hipMemcpyAsync(X, P->sb.X, bytes, out_kind, P->stream));
    HIPCHK(hipStreamSynchronize(P->stream));
    return MF_OK;
}
extern "C" int mf_solve_batch(mf_plan *P, int64_t nmat, int64_t nrhs, const double *B, double *X,
                              int32_t refine_steps, double bwd_tol, mf_solve_report *rep)
{
    return solve_batch(P, nmat, nrhs, B, X, refine_steps, bwd_tol, rep, hipMemcpyHostToDevice, hipMemcpyDeviceToHost);
}
extern "C" int mf_solve_batch_device(mf_plan *P, int64_t nmat, int64_t nrhs, const double *d_B, double *d_X,
                                     int32_t refine_steps, double bwd_tol, mf_solve_report *rep)
{
    return solve_batch(P, nmat, nrhs, d_B, d_X, refine_steps, bwd_tol, rep, hipMemcpyDeviceToDevice,
                       hipMemcpyDeviceToDevice);
}

// ------------------------------------------------------------------ export
static int export_factors(mf_plan *P, int64_t mat, int64_t *Lcolptr, int64_t *Lrowval, double *Lnzval,
                          int64_t *Ucolptr, int64_t *Urowval, double *Unzval, int64_t *p, int64_t *q)
{
    if (!P) return fail(MF_BADARG, "NULL plan");
    if (!P->factored) return fail(MF_BADARG, "plan has no factorization");
    if (mat < 0 || mat >= P->nmat_factored) return fail(MF_BADARG, "matrix index outside the factored batch");
    const int64_t n = P->n;
    const int32_t *d_prow = P->d_prow.as<int32_t>() + mat * n;
    if (Lcolptr)
        for (int64_t j = 0; j <= n; j++) Lcolptr[j] = P->Lp[j] + 1;
    if (Ucolptr)
        for (int64_t j = 0; j <= n; j++) Ucolptr[j] = P->Up[j] + 1;
    std::vector<int32_t> prow(n);
    HIPCHK(hipMemcpy(prow.data(), d_prow, (size_t)n * 4, hipMemcpyDeviceToHost));
    if (p)
        for (int64_t k = 0; k < n; k++) p[k] = P->qf[prow[k]] + 1;
    if (q)
        for (int64_t k = 0; k < n; k++) q[k] = P->qf[k] + 1;
    if (Lrowval || Lnzval || Urowval || Unzval) {
        HIPCHK(P->d_rowfac.alloc((size_t)n * 4));
        HIPCHK(upload(P->d_Lp, P->Lp));
        HIPCHK(upload(P->d_Up, P->Up));
        HIPCHK(P->d_Li.alloc((size_t)P->nnzL * 8));
        HIPCHK(P->d_Lx.alloc((size_t)P->nnzL * 8));
        HIPCHK(P->d_Ui.alloc((size_t)P->nnzU * 8));
        HIPCHK(P->d_Ux.alloc((size_t)P->nnzU * 8));
        HIPCHK(launch_rowfac(d_prow, P->d_rowfac.as<int32_t>(), n, P->stream));
        if (P->prec == MF_FP64)
            HIPCHK(launch_export<double>(P->d_sn.as<SnDev>(), P->nsuper, P->d_store.as<double>() + mat * P->store_elems,
                                         P->d_upd.as<int32_t>(),
                                         P->d_rowfac.as<int32_t>(), P->d_Lp.as<int64_t>(), P->d_Up.as<int64_t>(),
                                         P->d_uoff.as<int64_t>(), P->d_Li.as<int64_t>(), P->d_Lx.as<double>(),
                                         P->d_Ui.as<int64_t>(), P->d_Ux.as<double>(), P->stream));
        else
            HIPCHK(launch_export<float>(P->d_sn.as<SnDev>(), P->nsuper, P->d_store.as<float>() + mat * P->store_elems,
                                        P->d_upd.as<int32_t>(),
                                        P->d_rowfac.as<int32_t>(), P->d_Lp.as<int64_t>(), P->d_Up.as<int64_t>(),
                                        P->d_uoff.as<int64_t>(), P->d_Li.as<int64_t>(), P->d_Lx.as<double>(),
                                        P->d_Ui.as<int64_t>(), P->d_Ux.as<double>(), P->stream));
        HIPCHK(hipStreamSynchronize(P->stream));
        if (Lrowval) HIPCHK(hipMemcpy(Lrowval, P->d_Li.p, (size_t)P->nnzL * 8, hipMemcpyDeviceToHost));
        if (Lnzval) HIPCHK(hipMemcpy(Lnzval, P->d_Lx.p, (size_t)P->nnzL * 8, hipMemcpyDeviceToHost));
        if (Urowval) HIPCHK(hipMemcpy(Urowval, P->d_Ui.p, (size_t)P->nnzU * 8, hipMemcpyDeviceToHost));
        if (Unzval) HIPCHK(hipMemcpy(Unzval, P->d_Ux.p, (size_t)P->nnzU * 8, hipMemcpyDeviceToHost));
    }
    return MF_OK;
}

extern "C" int mf_export_factors(mf_plan *P, int64_t *Lcolptr, int64_t *Lrowval, double *Lnzval, int64_t *Ucolptr,
                                 int64_t *Urowval, double *Unzval, int64_t *p, int64_t *q)
{
    return export_factors(P, 0, Lcolptr, Lrowval, Lnzval, Ucolptr, Urowval, Unzval, p, q);
}

extern "C" int mf_export_factors_batch(mf_plan *P, int64_t mat, int64_t *Lcolptr, int64_t *Lrowval, double *Lnzval,
                                       int64_t *Ucolptr, int64_t *Urowval, double *Unzval, int64_t *p, int64_t *q)
{
    return export_factors(P, mat, Lcolptr, Lrowval, Lnzval, Ucolptr, Urowval, Unzval, p, q);
}

extern "C" int mf_plan_set_profiling(mf_plan *P, int32_t on)
{
    if (!P) return fail(MF_BADARG, "NULL plan");
    P->prof = on ? 1 : 0;
    for (int c = 0; c < MF_PROF_CLASSES; c++) {
        P->prof_ms[c] = P->prof_flops[c] = P->prof_bytes[c] = 0;
        P->prof_n[c] = 0;
    }
    P->evrec.clear();
    P->evused = 0;
    return MF_OK;
}

extern "C" int mf_plan_get_profile(const mf_plan *P, double *ms, int64_t *launches, double *flops, double *bytes)
{
    if (!P) return fail(MF_BADARG, "NULL plan");
    for (int c = 0; c < MF_PROF_CLASSES; c++) {
        if (ms) ms[c] = P->prof_ms[c];
        if (launches) launches[c] = P->prof_n[c];
        if (flops) flops[c] = P->prof_flops[c];
        if (bytes) bytes[c] = P->prof_bytes[c];
    }
    return MF_OK;
}

// ------------------------------------------------------------------ tile policy and timing
extern "C" int mf_set_gemm_tiles(int32_t precision, const int32_t *policy)
{
    if (!policy) return fail(MF_BADARG, "NULL policy");
    if (precision != MF_FP64 && precision != MF_FP32 && precision != MF_BF16) return fail(MF_BADARG, "bad precision");
    for (int i = 0; i < 2 * GEMM_NBUCKETS; i++)
        if (policy[i] < 0 || policy[i] >= GEMM_NTILES) return fail(MF_BADARG, "tile index out of range");
    std::memcpy(g_tile_policy[precision == MF_FP64 ? 0 : 1], policy, sizeof(int32_t) * 2 * GEMM_NBUCKETS);
    return MF_OK;
}

extern "C" int mf_get_gemm_tiles(int32_t precision, int32_t *policy)
{
    if (!policy) return fail(MF_BADARG, "NULL policy");
    std::memcpy(policy, g_tile_policy[precision == MF_FP64 ? 0 : 1], sizeof(int32_t) * 2 * GEMM_NBUCKETS);
    return MF_OK;
}

extern "C" int mf_gemm_tile_info(int32_t *ntiles, int32_t *tile_m, int32_t *tile_n, int32_t *nbuckets,
                                 int32_t *bucket_hi)
{
    if (ntiles) *ntiles = GEMM_NTILES;
    if (nbuckets) *nbuckets = GEMM_NBUCKETS;
    for (int t = 0; t < GEMM_NTILES; t++) {
        if (tile_m) tile_m[t] = GEMM_TILE_M[t];
        if (tile_n) tile_n[t] = GEMM_TILE_N[t];
    }
    if (bucket_hi)
        for (int b = 0; b + 1 < GEMM_NBUCKETS; b++) bucket_hi[b] = GEMM_BUCKET_HI[b];
    return MF_OK;
}

// Times nprob independent C(m x n) -= A(m x k) B(k x n) problems in one launch of tile shape
// `tile` (the autotuner's measurement): average seconds per launch over reps launches.
extern "C" int mf_time_gemm(int32_t precision, int64_t m, int64_t n, int64_t k, int64_t nprob, int32_t tile,
                            int32_t reps, double *seconds)
{
    if (!seconds || m < 1 || n < 1 || k < 1 || nprob < 1 || reps < 1) return fail(MF_BADARG, "bad gemm timing request");
    if (tile < 0 || tile >= GEMM_NTILES) return fail(MF_BADARG, "tile index out of range");
    if (m > INT32_MAX / 2 || n > INT32_MAX / 2 || k > INT32_MAX / 2) return fail(MF_BADARG, "gemm too large");
    const size_t esz = precision == MF_FP64 ? 8 : 4;
    const int64_t per = m * k + k * n + m * n;
    DevBuf pool, dprobs, dtiles;
    HIPCHK(pool.alloc((size_t)per * nprob * esz));
    HIPCHK(hipMemset(pool.p, 0, (size_t)per * nprob * esz));
    std::vector<GemmProb> pv;
    std::vector<GemmTile> tv;
    for (int64_t q = 0; q < nprob; q++) {
        GemmProb g{};
        g.aoff = q * per;
        g.boff = q * per + m * k;
        g.coff = q * per + m * k + k * n;
        g.m = (int32_t)m; g.n = (int32_t)n; g.k = (int32_t)k;
        g.lda = (int32_t)m; g.ldb = (int32_t)k; g.ldc = (int32_t)m;
        pv.push_back(g);
        for (int64_t tn = 0; tn < n; tn += GEMM_TILE_N[tile])
            for (int64_t tm = 0; tm < m; tm += GEMM_TILE_M[tile]) tv.push_back({(int32_t)q, (int32_t)tm, (int32_t)tn, 0});
    }
    xcd_order(tv, 0);
    HIPCHK(upload(dprobs, pv));
    HIPCHK(upload(dtiles, tv));
    hipStream_t st;
    HIPCHK(hipStreamCreate(&st));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    auto launch = [&]() {
        return esz == 8 ? launch_gemm_tile<double>(tile, 1, dprobs.as<GemmProb>(), dtiles.as<GemmTile>(), (int)tv.size(),
                                                    pool.p, 0, 1, nullptr, st)
                        : launch_gemm_tile<float>(tile, 1, dprobs.as<GemmProb>(), dtiles.as<GemmTile>(), (int)tv.size(),
                                                   pool.p, 0, 1, nullptr, st);
    };
    hipError_t err = launch();   // warm-up
    if (err == hipSuccess) err = hipEventRecord(e0, st);
    for (int r = 0; r < reps && err == hipSuccess; r++) err = launch();
    if (err == hipSuccess) err = hipEventRecord(e1, st);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    float ms = 0;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    HIPCHK(err);
    *seconds = ms * 1e-3 / reps;
    return MF_OK;
}

// Times one launch of a list of independent problems C(m x n) -= A(m x k) B(k x n) (mnk: nprob
// triples) with tile shape `tile`: the autotuner's measurement on a plan's real gemm launches.
extern "C" int mf_time_gemm_list(int32_t precision, int64_t nprob, const int32_t *mnk, int32_t tile, int32_t reps,
                                 double *seconds)
{
    if (!seconds || !mnk || nprob < 1 || reps < 1) return fail(MF_BADARG, "bad gemm timing request");
    if (tile < 0 || tile >= GEMM_NTILES) return fail(MF_BADARG, "tile index out of range");
    const size_t esz = precision == MF_FP64 ? 8 : 4;
    std::vector<GemmProb> pv;
    std::vector<GemmTile> tv;
    int64_t off = 0;
    for (int64_t q = 0; q < nprob; q++) {
        const int64_t m = mnk[3 * q], n = mnk[3 * q + 1], k = mnk[3 * q + 2];
        if (m < 1 || n < 1 || k < 1 || m > INT32_MAX / 2 || n > INT32_MAX / 2 || k > INT32_MAX / 2)
            return fail(MF_BADARG, "bad gemm problem size");
        GemmProb g{};
        g.aoff = off;
        g.boff = off + m * k;
        g.coff = off + m * k + k * n;
        g.m = (int32_t)m; g.n = (int32_t)n; g.k = (int32_t)k;
        g.lda = (int32_t)m; g.ldb = (int32_t)k; g.ldc = (int32_t)m;
        off += m * k + k * n + m * n;
        for (int64_t tn = 0; tn < n; tn += GEMM_TILE_N[tile])
            for (int64_t tm = 0; tm < m; tm += GEMM_TILE_M[tile]) tv.push_back({(int32_t)q, (int32_t)tm, (int32_t)tn, 0});
        pv.push_back(g);
    }
    {
        std::vector<int64_t> kw;
        for (const GemmProb &g : pv) kw.push_back(g.k);
        xcd_order_by_problem(tv, 0, kw);
    }
    DevBuf pool, dprobs, dtiles;
    HIPCHK(pool.alloc((size_t)off * esz));
    HIPCHK(hipMemset(pool.p, 0, (size_t)off * esz));
    HIPCHK(upload(dprobs, pv));
    HIPCHK(upload(dtiles, tv));
    hipStream_t st;
    HIPCHK(hipStreamCreate(&st));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    auto launch = [&]() {
        return esz == 8 ? launch_gemm_tile<double>(tile, 1, dprobs.as<GemmProb>(), dtiles.as<GemmTile>(), (int)tv.size(),
                                                    pool.p, 0, 1, nullptr, st)
                        : launch_gemm_tile<float>(tile, 1, dprobs.as<GemmProb>(), dtiles.as<GemmTile>(), (int)tv.size(),
                                                   pool.p, 0, 1, nullptr, st);
    };
    hipError_t err = launch();
    if (err == hipSuccess) err = hipEventRecord(e0, st);
    for (int r = 0; r < reps && err == hipSuccess; r++) err = launch();
    if (err == hipSuccess) err = hipEventRecord(e1, st);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    float ms = 0;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    HIPCHK(err);
    *seconds = ms * 1e-3 / reps;
    return MF_OK;
}

// A plain copy of `bytes` bytes between two fresh device buffers (16-byte vectors): the known
// traffic against which hardware byte counters are calibrated.
extern "C" int mf_copy_probe(int64_t bytes, double *seconds)
{
    if (bytes < 16 || bytes % 16) return fail(MF_BADARG, "copy probe size must be a positive multiple of 16");
    DevBuf src, dst;
    HIPCHK(src.alloc((size_t)bytes));
    HIPCHK(dst.alloc((size_t)bytes));
    HIPCHK(hipMemset(src.p, 1, (size_t)bytes));
    HIPCHK(hipDeviceSynchronize());
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    hipError_t err = hipEventRecord(e0, 0);
    if (err == hipSuccess) err = launch_copy(dst.p, src.p, bytes, 0);
    if (err == hipSuccess) err = hipEventRecord(e1, 0);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    float ms = 0;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCHK(err);
    if (seconds) *seconds = ms * 1e-3;
    return MF_OK;
}
