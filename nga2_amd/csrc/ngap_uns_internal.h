// Internal declarations of the unstructured solver (libngapuns.so).
#pragma once

#include <vector>

#include "ngapuns.h"

namespace ngap {
namespace amg {

// CSR with int32 indices (rows and nonzeros of every level < 2^31 are checked
// where the matrix is made)
struct Csr {
    int n = 0, ncol = 0;
    std::vector<int> rp, col;
    std::vector<double> val;
};

struct Hier {
    std::vector<Csr> A;               // levels 0..L-1
    std::vector<Csr> P;               // interpolation from level l+1 to l, l < L-1
    std::vector<std::vector<int>> cf; // C (1) / F (-1) split of level l, l < L-1
    std::vector<std::vector<double>> d; // l1 row norms of every level
    std::vector<double> Ainv;         // dense inverse of the coarsest level (row-major, symmetric)
    int singular = 0;
};

// The coarsest level is inverted as a dense matrix: at most this many rows.
constexpr int MAX_DENSE_ROWS = 4096;

Csr transpose(const Csr &A);
Csr spgemm(const Csr &A, const Csr &B);
int coarse_inverse(const Csr &A, std::vector<double> &X, int &singular);
int build(Hier &H, Csr A, const ngap_uns_opts &o);
// build() with host allocation failures, int32 index overflow and thread
// start failures returned as NGAP_ERR_ALLOC / NGAP_ERR_ARG / NGAP_ERR_ALLOC
int build_checked(Hier &H, Csr A, const ngap_uns_opts &o);
// the knobs' ranges (max_coarse <= MAX_DENSE_ROWS among them)
bool opts_valid(const ngap_uns_opts &o);

}  // namespace amg

long long umap(const ngap_grid &g, const double *opr, int nst, int *ind);
long long assemble(const ngap_grid &g, int nst, const int *stc, const int *ind, const double *opr, int *rowptr,
                   int *col, double *val);

}  // namespace ngap

struct ngap_amg {
    ngap::amg::Hier H;
};
