// mf_analysis.cpp -- host symbolic analysis for the multifrontal factor.
//
// In the drop-in the reference keeps this step in Julia (symbolic_mf, src/symbolic.jl:211-226)
// and passes its arrays to mf_plan_create.  This C++ restatement serves hosts without the
// reference (the Python mirror and the tests) and the faster analysis that a 10x quicker
// numeric phase needs.  It reproduces the reference's output exactly:
//   * AMD: the quotient-graph approximate minimum degree of CSparse cs_amd (Davis, "Direct
//     Methods for Sparse Linear Systems", SIAM 2006, sec. 7.1), as restated by
//     src/amd.jl:191-539 -- same pivot selection, element absorption, degree
//     approximation, supervariable hashing and final assembly-tree postorder, so ties break
//     identically;
//   * the A+A^T pattern (src/amd.jl:84-87) or the A^T A pattern with dense rows dropped
//     (src/amd.jl:91-104);
//   * etree (Liu, path compression), postorder with ascending children, the union-rule
//     column structure and the fundamental-supernode rule of src/symbolic.jl:67-194.
// Column structures are kept only while a column's parent still needs them, plus the
// structure of every supernode's last column (the update rows the numeric phase uses).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "mf_internal.h"

namespace mf {

using i64 = int64_t;

// 0-based CSC pattern with the diagonal removed and rows ascending.
struct Pattern {
    i64 n = 0;
    std::vector<i64> p, i;
};

static Pattern pattern_sym(i64 n, const i64 *Ap, const i64 *Ai)  // A + A^T
{
    // count A and A^T entries per column, then merge sorted unique rows
    std::vector<i64> tp(n + 1, 0);
    i64 nnz = Ap[n] - 1;
    for (i64 k = 0; k < nnz; k++) tp[Ai[k]]++;  // Ai 1-based: row r -> column r of A^T
    std::vector<i64> tcur(n + 1, 0);
    for (i64 j = 0; j < n; j++) tcur[j + 1] = tcur[j] + tp[j + 1];
    std::vector<i64> ti(std::max<i64>(nnz, 1));
    {
        std::vector<i64> cur(tcur.begin(), tcur.end() - 1);
        for (i64 j = 0; j < n; j++)
            for (i64 k = Ap[j] - 1; k < Ap[j + 1] - 1; k++) ti[cur[Ai[k] - 1]++] = j;  // sorted by j
    }
    Pattern S;
    S.n = n;
    S.p.assign(n + 1, 0);
    S.i.reserve(2 * nnz);
    std::vector<i64> colA;
    for (i64 j = 0; j < n; j++) {
        colA.clear();
        for (i64 k = Ap[j] - 1; k < Ap[j + 1] - 1; k++) colA.push_back(Ai[k] - 1);
        std::sort(colA.begin(), colA.end());
        // merge colA with ti[tcur[j]..tcur[j+1]) (already ascending)
        i64 a = 0, na = (i64)colA.size(), b = tcur[j], nb = tcur[j + 1];
        i64 last = -1;
        while (a < na || b < nb) {
            i64 r;
            if (b >= nb || (a < na && colA[a] <= ti[b])) r = colA[a++];
            else r = ti[b++];
            if (r == last) continue;
            last = r;
            if (r != j) S.i.push_back(r);
        }
        S.p[j + 1] = (i64)S.i.size();
    }
    return S;
}

static Pattern pattern_ata(i64 n, const i64 *Ap, const i64 *Ai, double dense_param)  // A^T A
{
    i64 m = n;  // square
    i64 nnz = Ap[n] - 1;
    // rows of A (A^T columns)
    std::vector<i64> rp(m + 1, 0);
    for (i64 k = 0; k < nnz; k++) rp[Ai[k]]++;
    for (i64 r = 0; r < m; r++) rp[r + 1] += rp[r];
    std::vector<i64> rj(std::max<i64>(nnz, 1));
    {
        std::vector<i64> cur(rp.begin(), rp.end() - 1);
        for (i64 j = 0; j < n; j++)
            for (i64 k = Ap[j] - 1; k < Ap[j + 1] - 1; k++) rj[cur[Ai[k] - 1]++] = j;
    }
    i64 thr = (i64)std::floor(std::min((double)(n - 2), std::max(16.0, dense_param * std::sqrt((double)n))));
    // (A^T A)(i, j) != 0 iff some kept row r holds both i and j
    Pattern S;
    S.n = n;
    S.p.assign(n + 1, 0);
    std::vector<i64> mark(n, -1), col;
    for (i64 j = 0; j < n; j++) {
        col.clear();
        for (i64 k = Ap[j] - 1; k < Ap[j + 1] - 1; k++) {
            i64 r = Ai[k] - 1;
            if (rp[r + 1] - rp[r] > thr) continue;
            for (i64 t = rp[r]; t < rp[r + 1]; t++) {
                i64 i = rj[t];
                if (i != j && mark[i] != j) { mark[i] = j; col.push_back(i); }
            }
        }
        std::sort(col.begin(), col.end());
        S.i.insert(S.i.end(), col.begin(), col.end());
        S.p[j + 1] = (i64)S.i.size();
    }
    return S;
}

// ---------------------------------------------------------------- AMD
namespace {
inline i64 flip(i64 i) { return -i - 2; }

struct Amd {
    i64 n;
    std::vector<i64> Ci, len, nv, next, head, elen, degree, w, hhead, last, Cp;

    i64 wclear(i64 mark, i64 lemax)
    {
        if (mark < 2 || mark + lemax < 0) {
            for (i64 k = 0; k < n; k++)
                if (w[k] != 0) w[k] = 1;
            mark = 2;
        }
        return mark;
    }

    static i64 tdfs(i64 j, i64 k, i64 *hd, const i64 *nx, i64 *post, i64 *stack)
    {
        i64 top = 0;
        stack[0] = j;
        while (top >= 0) {
            i64 p = stack[top];
            i64 i = hd[p];
            if (i == -1) {
                top--;
                post[k++] = p;
            } else {
                hd[p] = nx[i];
                stack[++top] = i;
            }
        }
        return k;
    }

    std::vector<i64> run(const Pattern &S, double dense)
    {
        n = S.n;
        i64 cnz = S.p[n];
        i64 nzmax = std::max<i64>(cnz + cnz / 5 + 2 * n, 1);
        Ci.assign(nzmax, 0);
        std::copy(S.i.begin(), S.i.end(), Ci.begin());
        double dthr = std::max(16.0, dense * std::sqrt((double)n));
        i64 densethr = std::min<i64>(n - 2, (i64)std::floor(dthr));
        len.assign(n + 1, 0);
        Cp.assign(n + 1, 0);
        for (i64 k = 0; k < n; k++) { len[k] = S.p[k + 1] - S.p[k]; Cp[k] = S.p[k]; }
        Cp[n] = S.p[n];
        nv.assign(n + 1, 1);
        next.assign(n + 1, -1);
        head.assign(n + 1, -1);
        elen.assign(n + 1, 0);
        degree = len;
        w.assign(n + 1, 1);
        hhead.assign(n + 1, -1);
        last.assign(n + 1, -1);
        i64 mark = wclear(0, 0);
        elen[n] = -2;
        Cp[n] = -1;
        w[n] = 0;
        i64 nel = 0;
        for (i64 i = 0; i < n; i++) {
            i64 d = degree[i];
            if (d == 0) {
                elen[i] = -2; nel++; Cp[i] = -1; w[i] = 0;
            } else if (d > densethr) {
                nv[i] = 0; elen[i] = -1; nel++; Cp[i] = flip(n); nv[n]++;
            } else {
                if (head[d] != -1) last[head[d]] = i;
                next[i] = head[d];
                head[d] = i;
            }
        }
        i64 mindeg = 0, lemax = 0;
        while (nel < n) {
            i64 k = -1;
            for (; mindeg < n; mindeg++) {
                k = head[mindeg];
                if (k != -1) break;
            }
            if (next[k] != -1) last[next[k]] = -1;
            head[mindeg] = next[k];
            i64 elenk = elen[k], nvk = nv[k];
            nel += nvk;

            if (elenk > 0 && cnz + mindeg >= nzmax) {  // compact Ci
                for (i64 j = 0; j < n; j++) {
                    i64 p = Cp[j];
                    if (p >= 0) { Cp[j] = Ci[p]; Ci[p] = flip(j); }
                }
                i64 q = 0;
                for (i64 p = 0; p < cnz;) {
                    i64 j = flip(Ci[p++]);
                    if (j >= 0) {
                        Ci[q] = Cp[j];
                        Cp[j] = q++;
                        for (i64 t = 0; t < len[j] - 1; t++) Ci[q++] = Ci[p++];
                    }
                }
                cnz = q;
            }

            // new element
            i64 dk = 0;
            nv[k] = -nvk;
            i64 p = Cp[k];
            i64 pk1 = (elenk == 0) ? p : cnz, pk2 = pk1;
            for (i64 k1 = 1; k1 <= elenk + 1; k1++) {
                i64 e, pj, ln;
                if (k1 > elenk) { e = k; pj = p; ln = len[k] - elenk; }
                else { e = Ci[p++]; pj = Cp[e]; ln = len[e]; }
                for (i64 k2 = 1; k2 <= ln; k2++) {
                    i64 i = Ci[pj++];
                    i64 nvi = nv[i];
                    if (nvi <= 0) continue;
                    dk += nvi;
                    nv[i] = -nvi;
                    Ci[pk2++] = i;
                    if (next[i] != -1) last[next[i]] = last[i];
                    if (last[i] != -1) next[last[i]] = next[i];
                    else head[degree[i]] = next[i];
                }
                if (e != k) { Cp[e] = flip(k); w[e] = 0; }
            }
            if (elenk != 0) cnz = pk2;
            degree[k] = dk;
            Cp[k] = pk1;
            len[k] = pk2 - pk1;
            elen[k] = -2;

            // set differences |Le \ Lk|
            mark = wclear(mark, lemax);
            for (i64 pk = pk1; pk < pk2; pk++) {
                i64 i = Ci[pk];
                i64 eln = elen[i];
                if (eln <= 0) continue;
                i64 nvi = -nv[i], wnvi = mark - nvi;
                for (i64 q = Cp[i]; q <= Cp[i] + eln - 1; q++) {
                    i64 e = Ci[q];
                    if (w[e] >= mark) w[e] -= nvi;
                    else if (w[e] != 0) w[e] = degree[e] + wnvi;
                }
            }

            // degree update and element absorption
            for (i64 pk = pk1; pk < pk2; pk++) {
                i64 i = Ci[pk];
                i64 p1 = Cp[i], p2 = p1 + elen[i] - 1, pn = p1;
                i64 h = 0, d = 0;
                for (i64 q = p1; q <= p2; q++) {
                    i64 e = Ci[q];
                    if (w[e] != 0) {
                        i64 dext = w[e] - mark;
                        if (dext > 0) { d += dext; Ci[pn++] = e; h += e; }
                        else { Cp[e] = flip(k); w[e] = 0; }
                    }
                }
                elen[i] = pn - p1 + 1;
                i64 p3 = pn, p4 = p1 + len[i];
                for (i64 q = p2 + 1; q < p4; q++) {
                    i64 j = Ci[q];
                    i64 nvj = nv[j];
                    if (nvj <= 0) continue;
                    d += nvj;
                    Ci[pn++] = j;
                    h += j;
                }
                if (d == 0) {
                    Cp[i] = flip(k);
                    i64 nvi = -nv[i];
                    dk -= nvi; nvk += nvi; nel += nvi;
                    nv[i] = 0; elen[i] = -1;
                } else {
                    degree[i] = std::min(degree[i], d);
                    Ci[pn] = Ci[p3];
                    Ci[p3] = Ci[p1];
                    Ci[p1] = k;
                    len[i] = pn - p1 + 1;
                    h = (h < 0 ? -h : h) % n;
                    next[i] = hhead[h];
                    hhead[h] = i;
                    last[i] = h;
                }
            }
            degree[k] = dk;
            lemax = std::max(lemax, dk);
            mark = wclear(mark + lemax, lemax);

            // supervariable detection
            for (i64 pk = pk1; pk < pk2; pk++) {
                i64 i = Ci[pk];
                if (nv[i] >= 0) continue;
                i64 h = last[i];
                i = hhead[h];
                hhead[h] = -1;
                for (; i != -1 && next[i] != -1; i = next[i], mark++) {
                    i64 ln = len[i], eln = elen[i];
                    for (i64 q = Cp[i] + 1; q <= Cp[i] + ln - 1; q++) w[Ci[q]] = mark;
                    i64 jlast = i;
                    for (i64 j = next[i]; j != -1;) {
                        bool ok = (len[j] == ln) && (elen[j] == eln);
                        for (i64 q = Cp[j] + 1; ok && q <= Cp[j] + ln - 1; q++)
                            if (w[Ci[q]] != mark) ok = false;
                        if (ok) {
                            Cp[j] = flip(i);
                            nv[i] += nv[j];
                            nv[j] = 0;
                            elen[j] = -1;
                            j = next[j];
                            next[jlast] = j;
                        } else {
                            jlast = j;
                            j = next[j];
                        }
                    }
                }
            }

            // finalise the new element
            p = pk1;
            for (i64 pk = pk1; pk < pk2; pk++) {
                i64 i = Ci[pk];
                i64 nvi = -nv[i];
                if (nvi <= 0) continue;
                nv[i] = nvi;
                i64 d = degree[i] + dk - nvi;
                d = std::min(d, n - nel - nvi);
                if (head[d] != -1) last[head[d]] = i;
                next[i] = head[d];
                last[i] = -1;
                head[d] = i;
                mindeg = std::min(mindeg, d);
                degree[i] = d;
                Ci[p++] = i;
            }
            nv[k] = nvk;
            if ((len[k] = p - pk1) == 0) { Cp[k] = -1; w[k] = 0; }
            if (elenk != 0) cnz = p;
        }

        // postorder the assembly tree
        for (i64 i = 0; i < n; i++) Cp[i] = flip(Cp[i]);
        for (i64 j = 0; j <= n; j++) head[j] = -1;
        for (i64 j = n; j >= 0; j--) {
            if (nv[j] > 0) continue;
            next[j] = head[Cp[j]];
            head[Cp[j]] = j;
        }
        for (i64 e = n; e >= 0; e--) {
            if (nv[e] <= 0) continue;
            if (Cp[e] != -1) { next[e] = head[Cp[e]]; head[Cp[e]] = e; }
        }
        std::vector<i64> P(n + 1);
        for (i64 k = 0, i = 0; i <= n; i++)
            if (Cp[i] == -1) k = tdfs(i, k, head.data(), next.data(), P.data(), w.data());
        P.resize(n);
        return P;
    }
};
}  // namespace

std::vector<i64> amd_order(const Pattern &S, double dense)
{
    if (S.n == 0) return {};
    if (S.n == 1) return {0};
    Amd a;
    return a.run(S, dense);
}

// ---------------------------------------------------------------- etree & friends
static std::vector<i64> etree(const Pattern &S)
{
    i64 n = S.n;
    std::vector<i64> parent(n, -1), anc(n, -1);
    for (i64 k = 0; k < n; k++)
        for (i64 p = S.p[k]; p < S.p[k + 1]; p++) {
            i64 i = S.i[p];
            while (i != -1 && i < k) {
                i64 inext = anc[i];
                anc[i] = k;
                if (inext == -1) parent[i] = k;
                i = inext;
            }
        }
    return parent;
}

static std::vector<i64> postorder(const std::vector<i64> &parent)
{
    i64 n = (i64)parent.size();
    std::vector<i64> head(n, -1), nextc(n, -1), post, stack;
    post.reserve(n);
    for (i64 j = n - 1; j >= 0; j--) {
        i64 p = parent[j];
        if (p != -1) { nextc[j] = head[p]; head[p] = j; }
    }
    for (i64 j = 0; j < n; j++) {
        if (parent[j] != -1) continue;
        stack.assign(1, j);
        while (!stack.empty()) {
            i64 p = stack.back();
            i64 i = head[p];
            if (i == -1) { stack.pop_back(); post.push_back(p); }
            else { head[p] = nextc[i]; stack.push_back(i); }
        }
    }
    return post;
}

static Pattern permute_sym(const Pattern &S, const std::vector<i64> &perm)
{
    i64 n = S.n;
    std::vector<i64> pinv(n);
    for (i64 j = 0; j < n; j++) pinv[perm[j]] = j;
    Pattern R;
    R.n = n;
    R.p.assign(n + 1, 0);
    R.i.resize(S.i.size());
    for (i64 j = 0; j < n; j++) R.p[j + 1] = R.p[j] + (S.p[perm[j] + 1] - S.p[perm[j]]);
    for (i64 j = 0; j < n; j++) {
        i64 o = perm[j], d = R.p[j];
        for (i64 p = S.p[o]; p < S.p[o + 1]; p++) R.i[d++] = pinv[S.i[p]];
        std::sort(R.i.begin() + R.p[j], R.i.begin() + R.p[j + 1]);
    }
    return R;
}

bool analyze(i64 n, const i64 *Ap, const i64 *Ai, int ordering, const i64 *q_in, Symbolic &out,
             std::string &err)
{
    out = Symbolic();
    out.n = n;
    std::vector<i64> q;
    if (ordering == 2) {
        if (!q_in) { err = "ordering MF_ORDER_GIVEN needs q"; return false; }
        q.resize(n);
        std::vector<char> seen(n, 0);
        for (i64 j = 0; j < n; j++) {
            i64 v = q_in[j] - 1;
            if (v < 0 || v >= n || seen[v]) { err = "q is not a permutation"; return false; }
            seen[v] = 1;
            q[j] = v;
        }
    } else if (ordering == 0) {
        q = amd_order(pattern_sym(n, Ap, Ai), 10.0);
    } else if (ordering == 1) {
        if (n <= 1) { q.resize(n); for (i64 j = 0; j < n; j++) q[j] = j; }
        else q = amd_order(pattern_ata(n, Ap, Ai, 10.0), 10.0);
    } else {
        err = "unknown ordering";
        return false;
    }
    // pattern of A[q,q] + A[q,q]^T = permutation of the A+A^T pattern
    Pattern S0 = pattern_sym(n, Ap, Ai);
    Pattern S1 = permute_sym(S0, q);
    std::vector<i64> parent1 = etree(S1);
    std::vector<i64> post = postorder(parent1);
    out.qf.resize(n);
    for (i64 j = 0; j < n; j++) out.qf[j] = q[post[j]];
    Pattern SF = permute_sym(S1, post);
    out.parent = etree(SF);
    const std::vector<i64> &parent = out.parent;

    // union-rule column structures, kept until the parent consumes them
    std::vector<i64> head(n, -1), nextc(n, -1), nchild(n, 0);
    for (i64 j = n - 1; j >= 0; j--) {
        i64 p = parent[j];
        if (p != -1) { nextc[j] = head[p]; head[p] = j; nchild[p]++; }
    }
    std::vector<std::vector<i64>> cs(n);
    std::vector<i64> colcount(n, 0), marker(n, -1);
    out.sstart.clear();
    out.ncolstruct = 0;
    std::vector<i64> super_of(n, 0);
    std::vector<i64> col;
    for (i64 j = 0; j < n; j++) {
        col.clear();
        marker[j] = j;
        for (i64 p = SF.p[j]; p < SF.p[j + 1]; p++) {
            i64 i = SF.i[p];
            if (i > j && marker[i] != j) { marker[i] = j; col.push_back(i); }
        }
        for (i64 c = head[j]; c != -1; c = nextc[c])
            for (i64 i : cs[c])
                if (i > j && marker[i] != j) { marker[i] = j; col.push_back(i); }
        std::sort(col.begin(), col.end());
        colcount[j] = (i64)col.size();
        out.ncolstruct += colcount[j];
        bool cont = j > 0 && parent[j - 1] == j && nchild[j] == 1 && colcount[j - 1] == colcount[j] + 1;
        if (!cont) {
            // j-1 closed a supernode: its structure is that supernode's update-row list.
            // cs[j-1] is still alive here because parent[j-1] >= j.
            if (j > 0) out.upd.push_back(cs[j - 1]);
            out.sstart.push_back(j);
        }
        super_of[j] = (i64)out.sstart.size() - 1;
        // every child of j has now been absorbed; nothing else reads its structure
        for (i64 c = head[j]; c != -1; c = nextc[c]) std::vector<i64>().swap(cs[c]);
        cs[j] = col;
    }
    if (n > 0) out.upd.push_back(cs[n - 1]);
    out.sstart.push_back(n);
    i64 ns = (i64)out.sstart.size() - 1;
    out.children.assign(ns, {});
    for (i64 s = 0; s < ns; s++) {
        i64 p = parent[out.sstart[s + 1] - 1];
        if (p != -1) out.children[super_of[p]].push_back(s);
    }
    return true;
}

// Relaxed supernode amalgamation (opt-in; it changes the L/U pattern, so the parity path never
// uses it).  Supernodes are visited in postorder; a supernode merges into its parent when its
// columns end right before the parent's (it is the child the postorder places last, so the
// merged column range stays contiguous) and either the merged pivot block has at most npmax
// columns or the explicit zeros the merge stores in L and U (the child's columns take the
// parent's rows) stay within a fraction zmax of the merged supernode's entries.  The child's
// own children become the merged supernode's.  The etree property (a child's update rows lie
// in its parent's front) keeps every original entry and every contribution block inside the
// merged fronts.  The result is again a SymbolicMF: supernode ranges, their update rows and
// children (ascending), the column etree of the filled structure and sum |colstruct[j]|.
int64_t relax_supernodes(Symbolic &S, double zmax, int64_t npmax)
{
    const i64 ns = (i64)S.sstart.size() - 1;
    if (ns <= 1) return 0;
    std::vector<i64> c1(ns), np(ns), nu(ns), zeros(ns, 0), par(ns, -1);
    std::vector<char> alive(ns, 1);
    for (i64 s = 0; s < ns; s++) {
        c1[s] = S.sstart[s];
        np[s] = S.sstart[s + 1] - S.sstart[s];
        nu[s] = (i64)S.upd[s].size();
        for (i64 c : S.children[s]) par[c] = s;
    }
    auto entries = [](i64 p, i64 u) { return p * (p + 1) + 2 * p * u; };   // L and U values of a supernode
    int64_t merged = 0;
    for (i64 s = 0; s < ns; s++) {
        const i64 p = par[s];
        if (p < 0 || S.sstart[s + 1] != c1[p]) continue;
        const i64 npm = np[s] + np[p];
        const i64 tot = entries(npm, nu[p]);
        const i64 z = tot - entries(np[s], nu[s]) - entries(np[p], nu[p]) + zeros[s] + zeros[p];
        if (npm > npmax && (double)z > zmax * (double)tot) continue;
        c1[p] = c1[s];
        np[p] = npm;
        zeros[p] = z;
        alive[s] = 0;
        merged++;
    }
    if (!merged) return 0;
    // new numbering of the surviving supernodes (still a postorder) and their children
    std::vector<i64> newid(ns, -1), root(ns);
    i64 m = 0;
    for (i64 s = 0; s < ns; s++)
        if (alive[s]) newid[s] = m++;
    for (i64 s = ns - 1; s >= 0; s--) root[s] = alive[s] ? s : root[par[s]];   // parents come later
    std::vector<i64> sstart(m + 1);
    std::vector<std::vector<i64>> upd(m), children(m);
    for (i64 s = 0; s < ns; s++) {
        if (!alive[s]) continue;
        const i64 t = newid[s];
        sstart[t] = c1[s];
        upd[t] = std::move(S.upd[s]);
    }
    sstart[m] = S.n;
    for (i64 s = 0; s < ns; s++)
        if (alive[s] && par[s] >= 0) children[newid[root[par[s]]]].push_back(newid[s]);
    for (auto &c : children) std::sort(c.begin(), c.end());
    S.sstart = std::move(sstart);
    S.upd = std::move(upd);
    S.children = std::move(children);
    // etree and column counts of the filled (relaxed) structure
    S.ncolstruct = 0;
    for (i64 t = 0; t < m; t++) {
        const i64 a = S.sstart[t], b = S.sstart[t + 1], w = b - a, u = (i64)S.upd[t].size();
        S.ncolstruct += w * (w - 1) / 2 + w * u;
        for (i64 j = a; j + 1 < b; j++) S.parent[j] = j + 1;
        S.parent[b - 1] = u > 0 ? S.upd[t][0] : -1;
    }
    return merged;
}

}  // namespace mf
