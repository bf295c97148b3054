// Host-only driver of the algebraic multigrid setup (nga2_amd/csrc/ngap_amg.cpp)
// for the AddressSanitizer / UndefinedBehaviorSanitizer build of
// tests/test_sanitizers.py: builds hierarchies of variable-coefficient 3-D
// Laplacians (periodic, walled, with a disconnected block) and reads every
// level back. Exit status 0 when every build succeeds.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "ngapuns.h"

static int laplacian_csr(int n, bool per, bool hole, std::vector<int> &rp, std::vector<int> &col,
                         std::vector<double> &val) {
    std::mt19937_64 rng(n * 131 + per);
    std::uniform_real_distribution<double> U(0.5, 2.0);
    const int N = n * n * n;
    auto id = [&](int i, int j, int k) { return i + n * (j + n * k); };
    auto inhole = [&](int i, int j, int k) { return hole && i >= n / 3 && i < n / 2 && j >= n / 3 && j < n / 2; };
    std::vector<double> w(3 * (size_t)N);
    for (auto &x : w) x = U(rng);
    rp.assign(N + 1, 0);
    col.clear();
    val.clear();
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                const int p = id(i, j, k);
                double diag = 0.0;
                std::vector<std::pair<int, double>> row;
                const int c[3] = {i, j, k};
                for (int d = 0; d < 3; d++)
                    for (int s = -1; s <= 1; s += 2) {
                        int q[3] = {i, j, k};
                        q[d] += s;
                        if (q[d] < 0 || q[d] >= n) {
                            if (!per) continue;
                            q[d] = (q[d] + n) % n;
                        }
                        const int pq = id(q[0], q[1], q[2]);
                        if (inhole(i, j, k) != inhole(q[0], q[1], q[2])) continue;  // the block is cut off
                        // the weight of the face between p and pq, stored at the upper cell of the pair
                        const double a = w[3 * (size_t)(s < 0 ? p : pq) + d];
                        (void)c;
                        row.push_back({pq, -a});
                        diag += a;
                    }
                row.push_back({p, diag + (per ? 0.0 : 1e-3)});
                std::sort(row.begin(), row.end());
                for (auto &e : row) {
                    col.push_back(e.first);
                    val.push_back(e.second);
                }
                rp[p + 1] = (int)col.size();
            }
    return N;
}

int main() {
    int fails = 0;
    for (int n : {6, 12, 20})
        for (int per = 0; per < 2; per++)
            for (int hole = 0; hole < 2; hole++) {
                std::vector<int> rp, col;
                std::vector<double> val;
                const int N = laplacian_csr(n, per, hole, rp, col, val);
                ngap_uns_opts o;
                ngap_uns_default_opts(&o);
                o.max_coarse = 20;
                ngap_amg *H = nullptr;
                const int rc = ngap_amg_build(&H, N, rp.data(), col.data(), val.data(), &o);
                if (rc) {
                    // a periodic system with a cut-off block has two constant null vectors; the
                    // coarsest dense inverse regularises one, so it may refuse (NGAP_ERR_COARSE)
                    std::printf("n=%d per=%d hole=%d: build rc %d\n", n, per, hole, rc);
                    if (!(rc == NGAP_ERR_COARSE && per && hole)) fails++;
                    continue;
                }
                const int L = ngap_amg_nlevels(H);
                for (int l = 0; l < L; l++) {
                    int m = 0, nnz = 0, pnnz = 0;
                    ngap_amg_level_sizes(H, l, &m, &nnz, &pnnz);
                    std::vector<int> arp(m + 1), acol(nnz), cf(m), prp(m + 1), pcol(pnnz > 0 ? pnnz : 1);
                    std::vector<double> aval(nnz), pval(pnnz > 0 ? pnnz : 1);
                    if (ngap_amg_level(H, l, arp.data(), acol.data(), aval.data(), cf.data(), prp.data(), pcol.data(),
                                       pval.data()))
                        fails++;
                }
                std::printf("n=%d per=%d hole=%d: %d levels\n", n, per, hole, L);
                ngap_amg_free(H);
            }
    return fails ? 1 : 0;
}
