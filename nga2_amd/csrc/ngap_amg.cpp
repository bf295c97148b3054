// Host side of the unstructured solver (libngapuns.so): the cell numbering
// and CSR assembly of NGA2's hypre_uns (hypre_uns_class.f90:186-230,
// :562-607) and the algebraic multigrid hierarchy hypre_uns asks BoomerAMG for
// (:27-33, :256-266), restated from the published algorithms:
//   strength     classical, theta (0.25), max row sum (0.9)
//   coarsening   PMIS (De Sterck, Yang, Heys 2006); the random part of the
//                measure is splitmix64 of the row index
//   interpolation extended+i (De Sterck, Falgout, Nolting, Yang 2008, eq. 4.9),
//                truncated to pmax (4) weights per row, positive and negative
//                weights rescaled separately to keep their sums
//   coarse operators Galerkin, A_c = P^T A P
//   coarsest     dense (pseudo-)inverse, gamma 1 1^T added when rows sum to 0
// Every sum is taken in a fixed order (row order of the operands), and
// contraction into fused multiply-adds is off, so the hierarchy is a
// deterministic function of the matrix.
#pragma clang fp contract(off)

#include "ngap_uns_internal.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <new>
#include <stdexcept>
#include <system_error>
#include <thread>
#include <utility>

namespace ngap {
namespace amg {

namespace {

constexpr int C_PT = 1, F_PT = -1, U_PT = 0;

// Row ranges on host threads (at most 16). Every row is computed by the same
// code whichever thread runs it, and per-thread outputs are joined in row
// order, so the hierarchy does not depend on the thread count.
int nthreads_for(long long n) {
    const unsigned hw = std::thread::hardware_concurrency();
    const long long cap = std::min<long long>(16, hw ? (long long)hw : 1);
    return (int)std::max<long long>(1, std::min<long long>(cap, n / 8192));
}

// threads for row-parallel work over a matrix of n rows and nnz nonzeros: the
// coarse levels have few rows but long ones (about 100 nonzeros per row three
// levels down), so their work follows the nonzeros
int nthreads_for(long long n, long long nnz) { return nthreads_for(std::max(n, nnz / 8)); }

// An exception in a worker (bad_alloc from a row's scratch vectors) is kept
// per thread and rethrown on the caller's thread after every started thread
// has been joined; a thread that cannot be started (system_error) stops the
// loop the same way. Either way no joinable std::thread is destroyed, so the
// caller's catch turns the failure into a return code instead of
// std::terminate.
template <typename F>
void par_rows(int n, int nt, F &&body) {
    if (nt <= 1) {
        body(0, n, 0);
        return;
    }
    std::vector<std::exception_ptr> err(nt);
    std::vector<std::thread> th;
    th.reserve(nt);  // emplace_back never reallocates below, so only the thread start can throw
    std::exception_ptr start_err;
    for (int t = 0; t < nt; t++) {
        const int lo = (int)((long long)n * t / nt), hi = (int)((long long)n * (t + 1) / nt);
        try {
            th.emplace_back([&body, &err, lo, hi, t] {
                try {
                    body(lo, hi, t);
                } catch (...) {
                    err[t] = std::current_exception();
                }
            });
        } catch (...) {
            start_err = std::current_exception();
            break;
        }
    }
    for (auto &x : th) x.join();
    if (start_err) std::rethrow_exception(start_err);
    for (auto &e : err)
        if (e) std::rethrow_exception(e);
}

// Row pointers from per-row counts (rp[i + 1] holds row i's count). Columns
// and values are indexed by int32, so a matrix of more nonzeros is refused.
void scan_counts(std::vector<int> &rp) {
    long long s = 0;
    for (size_t i = 1; i < rp.size(); i++) {
        s += rp[i];
        if (s > INT_MAX) throw std::length_error("nonzeros exceed int32 indexing");
        rp[i] = (int)s;
    }
}

double rand01(long long i) {
    uint64_t z = (uint64_t)i + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

double diag_of(const Csr &A, int i) {
    double d = 0.0;
    for (int q = A.rp[i]; q < A.rp[i + 1]; q++)
        if (A.col[q] == i) d += A.val[q];
    return d;
}

// S[q]: row i strongly depends on A.col[q]
std::vector<char> strength(const Csr &A, double theta, double max_row_sum) {
    std::vector<char> S(A.col.size(), 0);
    par_rows(A.n, nthreads_for(A.n, (long long)A.col.size()), [&](int lo, int hi, int) {
    for (int i = lo; i < hi; i++) {
        const int a = A.rp[i], b = A.rp[i + 1];
        const double diag = diag_of(A, i);
        double rsum = 0.0;
        for (int q = a; q < b; q++) rsum += A.val[q];
        if (max_row_sum < 1.0 && std::fabs(rsum) > std::fabs(diag) * max_row_sum) continue;
        bool any = false;
        double rs = 0.0;
        for (int q = a; q < b; q++) {
            if (A.col[q] == i) continue;
            if (!any) {
                rs = A.val[q];
                any = true;
            } else if (diag >= 0.0 ? A.val[q] < rs : A.val[q] > rs) {
                rs = A.val[q];
            }
        }
        if (!any) continue;
        if (diag >= 0.0) {
            if (rs >= 0.0) continue;
            const double thr = theta * rs;
            for (int q = a; q < b; q++)
                if (A.col[q] != i && A.val[q] < thr) S[q] = 1;
        } else {
            if (rs <= 0.0) continue;
            const double thr = theta * rs;
            for (int q = a; q < b; q++)
                if (A.col[q] != i && A.val[q] > thr) S[q] = 1;
        }
    }
    });
    return S;
}

// dep: i strongly depends on j (row order); inf: j depends on i (rows ascending)
struct Graph {
    std::vector<int> drp, dep, irp, inf;
};

// The dependencies row by row in parallel; the influences (the transpose)
// are counted and placed with atomic increments, then each list is sorted,
// which gives the ascending row order of a serial fill.
Graph strong_graph(const Csr &A, const std::vector<char> &S) {
    Graph G;
    const int n = A.n;
    const int nt = nthreads_for(n, (long long)A.col.size());
    G.drp.assign(n + 1, 0);
    G.irp.assign(n + 1, 0);
    par_rows(n, nt, [&](int lo, int hi, int) {
        for (int i = lo; i < hi; i++) {
            int cnt = 0;
            for (int q = A.rp[i]; q < A.rp[i + 1]; q++)
                if (S[q]) {
                    cnt++;
                    __atomic_fetch_add(&G.irp[A.col[q] + 1], 1, __ATOMIC_RELAXED);
                }
            G.drp[i + 1] = cnt;
        }
    });
    scan_counts(G.drp);
    scan_counts(G.irp);
    G.dep.resize(G.drp[n]);
    G.inf.resize(G.irp[n]);
    std::vector<int> pi(G.irp.begin(), G.irp.end() - 1);
    par_rows(n, nt, [&](int lo, int hi, int) {
        for (int i = lo; i < hi; i++) {
            int pd = G.drp[i];
            for (int q = A.rp[i]; q < A.rp[i + 1]; q++)
                if (S[q]) {
                    G.dep[pd++] = A.col[q];
                    G.inf[__atomic_fetch_add(&pi[A.col[q]], 1, __ATOMIC_RELAXED)] = i;
                }
        }
    });
    par_rows(n, nt, [&](int lo, int hi, int) {
        for (int j = lo; j < hi; j++) std::sort(G.inf.begin() + G.irp[j], G.inf.begin() + G.irp[j + 1]);
    });
    return G;
}

std::vector<int> pmis(const Csr &A, const Graph &G) {
    const int n = A.n;
    std::vector<double> w(n);
    std::vector<int> cf(n, U_PT);
    for (int i = 0; i < n; i++) {
        w[i] = (double)(G.irp[i + 1] - G.irp[i]) + rand01(i);
        if (w[i] < 1.0) cf[i] = F_PT;
    }
    std::vector<int> und, newc;
    std::vector<char> isnew(n, 0);
    for (;;) {
        und.clear();
        for (int i = 0; i < n; i++)
            if (cf[i] == U_PT) und.push_back(i);
        if (und.empty()) break;
        const int nu = (int)und.size(), nt = nthreads_for(nu);
        // new C points: undecided points beating every undecided strong neighbour
        // (reads the round's starting state only)
        std::vector<std::vector<int>> part(nt);
        par_rows(nu, nt, [&](int lo, int hi, int t) {
            for (int u = lo; u < hi; u++) {
                const int i = und[u];
                bool best = true;
                for (int q = G.drp[i]; q < G.drp[i + 1] && best; q++) {
                    const int j = G.dep[q];
                    if (cf[j] == U_PT && !(w[i] > w[j])) best = false;
                }
                for (int q = G.irp[i]; q < G.irp[i + 1] && best; q++) {
                    const int j = G.inf[q];
                    if (cf[j] == U_PT && !(w[i] > w[j])) best = false;
                }
                if (best) part[t].push_back(i);
            }
        });
        newc.clear();
        for (auto &pt : part) newc.insert(newc.end(), pt.begin(), pt.end());
        for (int i : newc) {
            cf[i] = C_PT;
            isnew[i] = 1;
        }
        // undecided points depending strongly on a new C point become F
        par_rows(nu, nt, [&](int lo, int hi, int) {
            for (int u = lo; u < hi; u++) {
                const int i = und[u];
                if (cf[i] != U_PT) continue;
                for (int q = G.drp[i]; q < G.drp[i + 1]; q++)
                    if (isnew[G.dep[q]]) {
                        cf[i] = F_PT;
                        break;
                    }
            }
        });
        for (int i : newc) isnew[i] = 0;
    }
    return cf;
}

// scratch of truncate_row, one per thread (no allocation per row)
struct TruncScratch {
    std::vector<int> key, keep, c2;
    std::vector<double> v2;
};

// keep the pmax largest |w| (ties: smaller coarse column first), rescale, sort by column
void truncate_row(std::vector<int> &cols, std::vector<double> &vals, int pmax, TruncScratch &w) {
    const int m = (int)cols.size();
    if (pmax > 0 && m > pmax) {
        w.key.resize(m);
        for (int q = 0; q < m; q++) w.key[q] = q;
        std::sort(w.key.begin(), w.key.end(), [&](int x, int y) {
            const double ax = std::fabs(vals[x]), ay = std::fabs(vals[y]);
            if (ax != ay) return ax > ay;
            return cols[x] < cols[y];
        });
        w.keep.assign(w.key.begin(), w.key.begin() + pmax);
        std::sort(w.keep.begin(), w.keep.end());
        double sp = 0.0, sn = 0.0, kp = 0.0, kn = 0.0;
        for (int q = 0; q < m; q++) {
            if (vals[q] > 0) sp += vals[q];
            if (vals[q] < 0) sn += vals[q];
        }
        for (int q : w.keep) {
            if (vals[q] > 0) kp += vals[q];
            if (vals[q] < 0) kn += vals[q];
        }
        const double fp = kp != 0.0 ? sp / kp : 1.0, fn = kn != 0.0 ? sn / kn : 1.0;
        w.c2.clear();
        w.v2.clear();
        for (int q : w.keep) {
            w.c2.push_back(cols[q]);
            w.v2.push_back(vals[q] * (vals[q] > 0 ? fp : fn));
        }
        cols.swap(w.c2);
        vals.swap(w.v2);
    }
    // stable insertion sort by column (rows are short; no temporary buffer)
    const int k = (int)cols.size();
    for (int u = 1; u < k; u++) {
        const int c = cols[u];
        const double v = vals[u];
        int q = u - 1;
        while (q >= 0 && cols[q] > c) {
            cols[q + 1] = cols[q];
            vals[q + 1] = vals[q];
            q--;
        }
        cols[q + 1] = c;
        vals[q + 1] = v;
    }
}

// rows [lo, hi) of P appended to (col, val), row lengths to len
void interp_rows(const Csr &A, const Graph &G, const std::vector<int> &cf, const std::vector<int> &cmap,
                 const std::vector<double> &adiag, int pmax, int lo, int hi, std::vector<int> &pcol,
                 std::vector<double> &pval, std::vector<int> &len) {
    const int n = A.n;
    // marks: in C-hat (position + 1), strong dependency of i (stamp)
    std::vector<int> chat_pos(n, 0), sstamp(n, -1);
    std::vector<int> chat, cols;
    std::vector<double> wv, vals;
    // entries of a strong F-neighbour's row that distribute to C-hat or to i,
    // in row order: (C-hat position, or -1 for i; value)
    std::vector<std::pair<int, double>> dk;
    TruncScratch tw;
    for (int i = lo; i < hi; i++) {
        const size_t before = pcol.size();
        if (cf[i] == C_PT) {
            pcol.push_back(cmap[i]);
            pval.push_back(1.0);
            len.push_back(1);
            continue;
        }
        chat.clear();
        for (int q = G.drp[i]; q < G.drp[i + 1]; q++) {
            const int j = G.dep[q];
            sstamp[j] = i;
            if (cf[j] == C_PT && !chat_pos[j]) {
                chat.push_back(j);
                chat_pos[j] = (int)chat.size();
            }
        }
        for (int q = G.drp[i]; q < G.drp[i + 1]; q++) {
            const int k = G.dep[q];
            if (cf[k] == C_PT) continue;
            for (int t = G.drp[k]; t < G.drp[k + 1]; t++) {
                const int j = G.dep[t];
                if (cf[j] == C_PT && !chat_pos[j]) {
                    chat.push_back(j);
                    chat_pos[j] = (int)chat.size();
                }
            }
        }
        if (!chat.empty()) {
            wv.assign(chat.size(), 0.0);
            double diag = 0.0;
            for (int q = A.rp[i]; q < A.rp[i + 1]; q++) {
                const int j = A.col[q];
                if (j == i)
                    diag += A.val[q];
                else if (chat_pos[j])
                    wv[chat_pos[j] - 1] += A.val[q];
                else if (sstamp[j] != i)
                    diag += A.val[q];  // weak, outside C-hat: lumped
            }
            for (int q = A.rp[i]; q < A.rp[i + 1]; q++) {
                const int k = A.col[q];
                if (k == i || chat_pos[k] || sstamp[k] != i) continue;
                const double aik = A.val[q];
                const double sgn = adiag[k] >= 0.0 ? 1.0 : -1.0;
                // one pass over row k: the denominator and the entries it distributes to
                double den = 0.0;
                dk.clear();
                for (int t = A.rp[k]; t < A.rp[k + 1]; t++) {
                    const int l = A.col[t];
                    const double v = A.val[t];
                    if (v * sgn < 0.0) {
                        if (chat_pos[l]) {
                            dk.emplace_back(chat_pos[l] - 1, v);
                            den += v;
                        } else if (l == i) {
                            dk.emplace_back(-1, v);
                            den += v;
                        }
                    }
                }
                if (den == 0.0) {
                    diag += aik;
                    continue;
                }
                const double dist = aik / den;
                for (const auto &e : dk) {
                    if (e.first >= 0)
                        wv[e.first] += dist * e.second;
                    else
                        diag += dist * e.second;
                }
            }
            cols.resize(chat.size());
            vals.resize(chat.size());
            for (size_t q = 0; q < chat.size(); q++) {
                cols[q] = cmap[chat[q]];
                vals[q] = -wv[q] / diag;
            }
            truncate_row(cols, vals, pmax, tw);
            pcol.insert(pcol.end(), cols.begin(), cols.end());
            pval.insert(pval.end(), vals.begin(), vals.end());
        }
        for (int j : chat) chat_pos[j] = 0;
        len.push_back((int)(pcol.size() - before));
    }
}

Csr interp_ext_i(const Csr &A, const Graph &G, const std::vector<int> &cf, int pmax) {
    const int n = A.n;
    std::vector<int> cmap(n, -1);
    int nc = 0;
    for (int i = 0; i < n; i++)
        if (cf[i] == C_PT) cmap[i] = nc++;
    std::vector<double> adiag(n);
    const int nt = nthreads_for(n, (long long)A.col.size());
    par_rows(n, nt, [&](int lo, int hi, int) {
        for (int i = lo; i < hi; i++) adiag[i] = diag_of(A, i);
    });
    std::vector<std::vector<int>> pcol(nt), len(nt);
    std::vector<std::vector<double>> pval(nt);
    par_rows(n, nt, [&](int lo, int hi, int t) {
        interp_rows(A, G, cf, cmap, adiag, pmax, lo, hi, pcol[t], pval[t], len[t]);
    });
    Csr P;
    P.n = n;
    P.ncol = nc;
    P.rp.assign(n + 1, 0);
    int row = 0;
    for (int t = 0; t < nt; t++)
        for (int l : len[t]) P.rp[++row] = l;
    scan_counts(P.rp);
    for (int t = 0; t < nt; t++) {
        P.col.insert(P.col.end(), pcol[t].begin(), pcol[t].end());
        P.val.insert(P.val.end(), pval[t].begin(), pval[t].end());
    }
    return P;
}

}  // namespace

// Rows of the transpose counted and filled with atomic increments in
// parallel, then each row sorted by column: the order of a serial fill.
Csr transpose(const Csr &A) {
    Csr T;
    T.n = A.ncol;
    T.ncol = A.n;
    T.rp.assign(T.n + 1, 0);
    const int nt = nthreads_for(A.n, (long long)A.col.size());
    par_rows(A.n, nt, [&](int lo, int hi, int) {
        for (int q = A.rp[lo]; q < A.rp[hi]; q++) __atomic_fetch_add(&T.rp[A.col[q] + 1], 1, __ATOMIC_RELAXED);
    });
    scan_counts(T.rp);
    T.col.resize(A.col.size());
    T.val.resize(A.col.size());
    std::vector<int> pos(T.rp.begin(), T.rp.end() - 1);
    par_rows(A.n, nt, [&](int lo, int hi, int) {
        for (int i = lo; i < hi; i++)
            for (int q = A.rp[i]; q < A.rp[i + 1]; q++) {
                const int p = __atomic_fetch_add(&pos[A.col[q]], 1, __ATOMIC_RELAXED);
                T.col[p] = i;
                T.val[p] = A.val[q];
            }
    });
    par_rows(T.n, nthreads_for(T.n, (long long)T.col.size()), [&](int lo, int hi, int) {
        for (int r = lo; r < hi; r++) {
            // insertion sort of the row's (column, value) pairs: rows are short
            const int a = T.rp[r], b = T.rp[r + 1];
            for (int u = a + 1; u < b; u++) {
                const int c = T.col[u];
                const double v = T.val[u];
                int w = u - 1;
                while (w >= a && T.col[w] > c) {
                    T.col[w + 1] = T.col[w];
                    T.val[w + 1] = T.val[w];
                    w--;
                }
                T.col[w + 1] = c;
                T.val[w + 1] = v;
            }
        }
    });
    return T;
}

// C = A B, columns of each row ascending. Two passes over the rows: the
// first counts each row's distinct columns, the second writes columns and
// sums in place. A row with up to HCAP/2 products accumulates in a small
// open-addressing table that stays in cache, a longer one in a dense array.
// Each column sums its products in the order of A's row, then B's row.
namespace {
constexpr int HCAP = 4096;

struct RowAcc {
    std::vector<int> hkey, cols, slots, mark;
    std::vector<double> hval, acc;
    RowAcc() : hkey(HCAP, -1), hval(HCAP) {}

    // the distinct columns of row i of A B in cols; with vals set, ascending
    // and with their sums (read back through column())
    void row(const Csr &A, const Csr &B, int i, bool vals) {
        cols.clear();
        long long bound = 0;
        for (int q = A.rp[i]; q < A.rp[i + 1]; q++) bound += B.rp[A.col[q] + 1] - B.rp[A.col[q]];
        dense = bound > HCAP / 2;
        if (!dense) {
            slots.clear();
            for (int q = A.rp[i]; q < A.rp[i + 1]; q++) {
                const int k = A.col[q];
                const double a = A.val[q];
                for (int u = B.rp[k]; u < B.rp[k + 1]; u++) {
                    const int j = B.col[u];
                    unsigned h = ((unsigned)j * 2654435761u) & (HCAP - 1);
                    while (hkey[h] != -1 && hkey[h] != j) h = (h + 1) & (HCAP - 1);
                    if (hkey[h] == -1) {
                        hkey[h] = j;
                        hval[h] = 0.0;
                        slots.push_back((int)h);
                        cols.push_back(j);
                    }
                    if (vals) hval[h] += a * B.val[u];
                }
            }
        } else {
            if (acc.empty()) {
                acc.assign(B.ncol, 0.0);
                mark.assign(B.ncol, -1);
            }
            for (int q = A.rp[i]; q < A.rp[i + 1]; q++) {
                const int k = A.col[q];
                const double a = A.val[q];
                for (int u = B.rp[k]; u < B.rp[k + 1]; u++) {
                    const int j = B.col[u];
                    if (mark[j] != i) {
                        mark[j] = i;
                        acc[j] = 0.0;
                        cols.push_back(j);
                    }
                    if (vals) acc[j] += a * B.val[u];
                }
            }
            // a later row reuses mark: clear this row's entries
            for (int j : cols) mark[j] = -1;
        }
        // the counting pass needs the number of columns only
        if (vals) std::sort(cols.begin(), cols.end());
    }
    double column(int j) const {
        if (dense) return acc[j];
        unsigned h = ((unsigned)j * 2654435761u) & (HCAP - 1);
        while (hkey[h] != j) h = (h + 1) & (HCAP - 1);
        return hval[h];
    }
    void clear() {
        if (!dense)
            for (int h : slots) hkey[h] = -1;
    }
    bool dense = false;
};
}  // namespace

Csr spgemm(const Csr &A, const Csr &B) {
    const int nt = nthreads_for(A.n, (long long)A.col.size() + (long long)B.col.size());
    Csr C;
    C.n = A.n;
    C.ncol = B.ncol;
    C.rp.assign(A.n + 1, 0);
    par_rows(A.n, nt, [&](int lo, int hi, int) {
        RowAcc r;
        for (int i = lo; i < hi; i++) {
            r.row(A, B, i, false);
            C.rp[i + 1] = (int)r.cols.size();
            r.clear();
        }
    });
    scan_counts(C.rp);
    C.col.resize(C.rp[A.n]);
    C.val.resize(C.rp[A.n]);
    par_rows(A.n, nt, [&](int lo, int hi, int) {
        RowAcc r;
        for (int i = lo; i < hi; i++) {
            r.row(A, B, i, true);
            int p = C.rp[i];
            for (int j : r.cols) {
                C.col[p] = j;
                C.val[p] = r.column(j);
                p++;
            }
            r.clear();
        }
    });
    return C;
}

int coarse_inverse(const Csr &A, std::vector<double> &X, int &singular) {
    const int n = A.n;
    if (n > MAX_DENSE_ROWS) return NGAP_ERR_COARSE;  // coarsening stopped above the dense limit
    std::vector<double> M((size_t)n * n, 0.0);
    double rs = 0.0, dsum = 0.0;
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int q = A.rp[i]; q < A.rp[i + 1]; q++) {
            M[(size_t)i * n + A.col[q]] += A.val[q];
            s += A.val[q];
            if (A.col[q] == i) dsum += std::fabs(A.val[q]);
        }
        rs += std::fabs(s);
    }
    singular = rs <= 1e-10 * dsum;
    if (singular) {
        const double gam = dsum / ((double)n * n);
        for (auto &m : M) m += gam;
    }
    // Cholesky (the operator is symmetric positive definite, or semi-definite
    // made definite above); a non-positive pivot is reported
    std::vector<double> L((size_t)n * n, 0.0);
    for (int j = 0; j < n; j++) {
        double s = M[(size_t)j * n + j];
        for (int k = 0; k < j; k++) s -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
        if (!(s > 0.0)) return NGAP_ERR_COARSE;
        const double ljj = std::sqrt(s);
        L[(size_t)j * n + j] = ljj;
        for (int i = j + 1; i < n; i++) {
            double t = M[(size_t)i * n + j];
            for (int k = 0; k < j; k++) t -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
            L[(size_t)i * n + j] = t / ljj;
        }
    }
    X.assign((size_t)n * n, 0.0);
    std::vector<double> y(n);
    for (int col = 0; col < n; col++) {
        for (int i = 0; i < n; i++) {
            double s = (i == col) ? 1.0 : 0.0;
            for (int k = 0; k < i; k++) s -= L[(size_t)i * n + k] * y[k];
            y[i] = s / L[(size_t)i * n + i];
        }
        for (int i = n - 1; i >= 0; i--) {
            double s = y[i];
            for (int k = i + 1; k < n; k++) s -= L[(size_t)k * n + i] * y[k];
            y[i] = s / L[(size_t)i * n + i];
        }
        for (int i = 0; i < n; i++) X[(size_t)col * n + i] = y[i];
    }
    return 0;
}

int build(Hier &H, Csr A, const ngap_uns_opts &o) {
    H.A.clear();
    H.P.clear();
    H.cf.clear();
    H.A.push_back(std::move(A));
    while (H.A.back().n > o.max_coarse && (int)H.A.size() < o.max_levels) {
        const Csr &F = H.A.back();
        std::vector<char> S = strength(F, o.strong_threshold, o.max_row_sum);
        Graph G = strong_graph(F, S);
        std::vector<int> cf = pmis(F, G);
        int nc = 0;
        for (int c : cf) nc += c == C_PT;
        if (nc == 0 || nc >= F.n) break;
        Csr P = interp_ext_i(F, G, cf, o.pmax);
        Csr R = transpose(P);
        Csr AP = spgemm(F, P);
        Csr Ac = spgemm(R, AP);
        H.cf.push_back(std::move(cf));
        H.P.push_back(std::move(P));
        H.A.push_back(std::move(Ac));
    }
    H.d.clear();
    for (const Csr &L : H.A) {
        std::vector<double> d(L.n);
        par_rows(L.n, nthreads_for(L.n, (long long)L.col.size()), [&](int lo, int hi, int) {
            for (int i = lo; i < hi; i++) {
                double s = 0.0;
                for (int q = L.rp[i]; q < L.rp[i + 1]; q++) s += std::fabs(L.val[q]);
                d[i] = s;
            }
        });
        H.d.push_back(std::move(d));
    }
    return coarse_inverse(H.A.back(), H.Ainv, H.singular);
}

int build_checked(Hier &H, Csr A, const ngap_uns_opts &o) {
    try {
        return build(H, std::move(A), o);
    } catch (const std::bad_alloc &) {
        return NGAP_ERR_ALLOC;
    } catch (const std::length_error &) {
        return NGAP_ERR_ARG;
    } catch (const std::system_error &) {
        return NGAP_ERR_ALLOC;  // no host thread could be started
    }
}

bool opts_valid(const ngap_uns_opts &o) {
    return o.nu >= 1 && o.kdim >= 1 && o.kdim <= 30 && o.max_coarse >= 1 && o.max_coarse <= MAX_DENSE_ROWS &&
           o.max_levels >= 1 && o.pmax >= 0 && o.strong_threshold >= 0.0 && o.strong_threshold <= 1.0 &&
           o.max_row_sum > 0.0 && (o.precision == NGAP_FP64 || o.precision == NGAP_MIXED);
}

}  // namespace amg

// ------------------------------------------------------------------ numbering and matrix
long long umap(const ngap_grid &g, const double *opr, int nst, int *ind) {
    const long long nxo = g.imaxo_ - g.imino_ + 1, nyo = g.jmaxo_ - g.jmino_ + 1, nzo = g.kmaxo_ - g.kmino_ + 1;
    const long long ntot = nxo * nyo * nzo;
    for (long long q = 0; q < ntot; q++) ind[q] = -1;
    auto lin = [&](int i, int j, int k) {
        return (long long)(i - g.imino_) + nxo * ((long long)(j - g.jmino_) + nyo * (long long)(k - g.kmino_));
    };
    const double tiny = 10.0 * 2.220446049250313e-16;
    long long count = 0;
    for (int k = g.kmin_; k <= g.kmax_; k++)
        for (int j = g.jmin_; j <= g.jmax_; j++)
            for (int i = g.imin_; i <= g.imax_; i++) {
                const long long p = lin(i, j, k);
                if (std::fabs(opr[(size_t)nst * p]) > tiny) ind[p] = (int)count++;
            }
    // periodic images into the ghosts, direction by direction (cfg%sync)
    const int nx = g.imax_ - g.imin_ + 1, ny = g.jmax_ - g.jmin_ + 1, nz = g.kmax_ - g.kmin_ + 1;
    auto wrap = [](int v, int lo, int n) { return lo + (((v - lo) % n) + n) % n; };
    if (g.xper)
        for (int k = g.kmino_; k <= g.kmaxo_; k++)
            for (int j = g.jmino_; j <= g.jmaxo_; j++)
                for (int i = g.imino_; i <= g.imaxo_; i++)
                    if (i < g.imin_ || i > g.imax_) ind[lin(i, j, k)] = ind[lin(wrap(i, g.imin_, nx), j, k)];
    if (g.yper)
        for (int k = g.kmino_; k <= g.kmaxo_; k++)
            for (int j = g.jmino_; j <= g.jmaxo_; j++)
                if (j < g.jmin_ || j > g.jmax_)
                    for (int i = g.imino_; i <= g.imaxo_; i++) ind[lin(i, j, k)] = ind[lin(i, wrap(j, g.jmin_, ny), k)];
    if (g.zper)
        for (int k = g.kmino_; k <= g.kmaxo_; k++)
            if (k < g.kmin_ || k > g.kmax_)
                for (int j = g.jmino_; j <= g.jmaxo_; j++)
                    for (int i = g.imino_; i <= g.imaxo_; i++) ind[lin(i, j, k)] = ind[lin(i, j, wrap(k, g.kmin_, nz))];
    return count;
}

long long assemble(const ngap_grid &g, int nst, const int *stc, const int *ind, const double *opr, int *rowptr,
                   int *col, double *val) {
    const long long nxo = g.imaxo_ - g.imino_ + 1, nyo = g.jmaxo_ - g.jmino_ + 1;
    auto lin = [&](int i, int j, int k) {
        return (long long)(i - g.imino_) + nxo * ((long long)(j - g.jmino_) + nyo * (long long)(k - g.kmino_));
    };
    long long nnz = 0, row = 0;
    int rc[64];
    double rv[64];
    if (rowptr) rowptr[0] = 0;
    for (int k = g.kmin_; k <= g.kmax_; k++)
        for (int j = g.jmin_; j <= g.jmax_; j++)
            for (int i = g.imin_; i <= g.imax_; i++) {
                const long long p = lin(i, j, k);
                if (ind[p] < 0) continue;
                int m = 0;
                for (int st = 0; st < nst; st++) {
                    const int c = ind[lin(i + stc[st], j + stc[nst + st], k + stc[2 * nst + st])];
                    if (c < 0) continue;
                    const double v = opr[(size_t)nst * p + st];
                    int q = 0;
                    while (q < m && rc[q] != c) q++;
                    if (q < m) {
                        rv[q] += v;
                    } else {
                        rc[m] = c;
                        rv[m] = v;
                        m++;
                    }
                }
                if (col && val)
                    for (int q = 0; q < m; q++) {
                        col[nnz + q] = rc[q];
                        val[nnz + q] = rv[q];
                    }
                nnz += m;
                row++;
                if (rowptr) rowptr[row] = (int)nnz;
            }
    return nnz;
}

}  // namespace ngap

// ------------------------------------------------------------------ C ABI (host)
using namespace ngap;

extern "C" {

int ngap_uns_default_opts(ngap_uns_opts *o) {
    if (!o) return NGAP_ERR_ARG;
    o->strong_threshold = 0.25;
    o->max_row_sum = 0.9;
    o->pmax = 4;
    o->max_coarse = 512;
    o->max_levels = 25;
    o->nu = 1;
    o->kdim = 5;
    o->graphs = 1;
    o->verbose = 0;
    o->precision = NGAP_FP64;
    return 0;
}

long long ngap_uns_umap_host(const ngap_grid *g, const double *opr, int nst, int *ind) {
    if (!g || !opr || !ind || nst < 1) return NGAP_ERR_ARG;
    return umap(*g, opr, nst, ind);
}

long long ngap_uns_assemble_host(const ngap_grid *g, int nst, const int *stc, const int *ind, const double *opr,
                                 int *rowptr, int *col, double *val) {
    if (!g || !stc || !ind || !opr || nst < 1 || nst > 64) return NGAP_ERR_ARG;
    return assemble(*g, nst, stc, ind, opr, rowptr, col, val);
}

int ngap_amg_build(ngap_amg **H, int n, const int *rowptr, const int *col, const double *val,
                   const ngap_uns_opts *o) {
    if (!H || n < 1 || !rowptr || !col || !val) return NGAP_ERR_ARG;
    ngap_uns_opts d;
    ngap_uns_default_opts(&d);
    const ngap_uns_opts &op = o ? *o : d;
    if (!amg::opts_valid(op) || rowptr[0] != 0) return NGAP_ERR_ARG;
    for (int i = 0; i < n; i++)
        if (rowptr[i + 1] < rowptr[i]) return NGAP_ERR_ARG;
    for (long long q = 0; q < rowptr[n]; q++)
        if (col[q] < 0 || col[q] >= n) return NGAP_ERR_ARG;
    amg::Csr A;
    auto *h = new (std::nothrow) ngap_amg;
    if (!h) return NGAP_ERR_ALLOC;
    try {
        A.n = n;
        A.ncol = n;
        A.rp.assign(rowptr, rowptr + n + 1);
        A.col.assign(col, col + rowptr[n]);
        A.val.assign(val, val + rowptr[n]);
    } catch (const std::bad_alloc &) {
        delete h;
        return NGAP_ERR_ALLOC;
    }
    const int rc = amg::build_checked(h->H, std::move(A), op);
    if (rc) {
        delete h;
        return rc;
    }
    *H = h;
    return 0;
}

int ngap_amg_nlevels(const ngap_amg *H) { return H ? (int)H->H.A.size() : NGAP_ERR_ARG; }

int ngap_amg_level_sizes(const ngap_amg *H, int l, int *n, int *nnz, int *pnnz) {
    if (!H || l < 0 || l >= (int)H->H.A.size()) return NGAP_ERR_ARG;
    const amg::Csr &A = H->H.A[l];
    if (n) *n = A.n;
    if (nnz) *nnz = (int)A.col.size();
    if (pnnz) *pnnz = l < (int)H->H.P.size() ? (int)H->H.P[l].col.size() : 0;
    return 0;
}

int ngap_amg_level(const ngap_amg *H, int l, int *rowptr, int *col, double *val, int *cf, int *prowptr, int *pcol,
                   double *pval) {
    if (!H || l < 0 || l >= (int)H->H.A.size()) return NGAP_ERR_ARG;
    const amg::Csr &A = H->H.A[l];
    if (rowptr) std::memcpy(rowptr, A.rp.data(), sizeof(int) * (A.n + 1));
    if (col) std::memcpy(col, A.col.data(), sizeof(int) * A.col.size());
    if (val) std::memcpy(val, A.val.data(), sizeof(double) * A.val.size());
    if (l < (int)H->H.P.size()) {
        const amg::Csr &P = H->H.P[l];
        if (cf) std::memcpy(cf, H->H.cf[l].data(), sizeof(int) * A.n);
        if (prowptr) std::memcpy(prowptr, P.rp.data(), sizeof(int) * (P.n + 1));
        if (pcol) std::memcpy(pcol, P.col.data(), sizeof(int) * P.col.size());
        if (pval) std::memcpy(pval, P.val.data(), sizeof(double) * P.val.size());
    }
    return 0;
}

void ngap_amg_free(ngap_amg *H) { delete H; }

}  // extern "C"
