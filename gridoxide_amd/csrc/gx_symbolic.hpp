// Host-side symbolic phase of the sparse LU used by the Newton step: block triangular form,
// minimum-degree ordering per block, and the first numeric factorization that fixes the pivot order.
// Runs once per topology; the device only ever refactors with the order found here.
#pragma once
#include <cstdint>
#include <vector>

namespace gx {

// Square sparse pattern in compressed-column form, rows sorted and unique within each column.
struct CscPattern {
  int n = 0;
  std::vector<int> ap;  // n+1
  std::vector<int> ai;  // nnz
};

// Result of the ordering phase (what KLU's analyze produces): the matrix is factored as
// A(p, q) with diagonal blocks [r[b], r[b+1]).
struct Ordering {
  std::vector<int> p, q, r;
};

// Maximum transversal: match[i] = column matched to row i (-1 if unmatched).  Columns are
// visited in order; each first tries an unmatched row in its own column (the cheap assignment,
// rows scanned in stored order), then searches augmenting paths depth first.  Returns the size
// of the matching.
int max_transversal(const CscPattern& a, std::vector<int>& match);

// Strongly connected components of the graph of A(:, colperm): node j has an edge to row i
// for every entry A(i, colperm[j]).  Roots are taken in increasing order and edges in stored
// order; blocks are numbered in the order their components complete (which makes the permuted
// matrix upper block triangular), and within a block nodes keep increasing index order.
// Writes p (nodes in block order) and r (block boundaries); returns the number of blocks.
int strong_components(const CscPattern& a, const std::vector<int>& colperm, std::vector<int>& p,
                      std::vector<int>& r);

// Approximate minimum degree ordering of the pattern of A + A' (diagonal ignored).
// Returns perm with perm[k] = the k-th variable to eliminate.
std::vector<int> amd_order(const CscPattern& a);

// The full ordering phase: BTF (maximum transversal + strong components), then per block AMD
// when the block has more than 3 columns, natural order otherwise.  Returns false when the
// matrix is structurally singular.
bool analyze(const CscPattern& a, Ordering& ord);

// LU factors of A(pnum, q) in pivot-step space, ready for refactoring:
//   column k holds U rows (j < k, ascending), then the diagonal, then L rows (i > k, ascending).
struct LuPattern {
  int n = 0;
  std::vector<int> q;       // pivot step k factors column q[k]
  std::vector<int> pnum;    // pivot step k takes row pnum[k]
  std::vector<int> pinv;    // row -> pivot step
  std::vector<int> cp;      // n+1 column starts in the packed L+U value array
  std::vector<int> ulen;    // number of strictly-upper entries of column k
  std::vector<int> row;     // pivot-step row index of every packed entry
  int nnz() const { return cp.empty() ? 0 : cp[n]; }
};

// First numeric factorization with partial pivoting (KLU's rule: the largest candidate unless
// the preferred diagonal row's candidate is at least tol times as large).  The preferred row of
// step k is ord.p[k].  Returns false when a column has no candidate or a zero pivot.
bool factor_pivot(const CscPattern& a, const double* ax, const Ordering& ord, double tol, LuPattern& lu);

// LU pattern for an externally supplied pivot order (no numeric pivoting).
bool pattern_from_order(const CscPattern& a, const std::vector<int>& q, const std::vector<int>& pnum,
                        LuPattern& lu);

}  // namespace gx
