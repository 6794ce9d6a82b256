// Symbolic phase of the Newton step's sparse LU (host, once per topology).
//
// Behaviour follows the reference's KluNative / SuiteSparse KLU configuration (BTF pre-ordering,
// AMD on each diagonal block larger than 3, partial pivoting with diagonal preference at
// tol = 1e-3, no row scaling): src/klu_native/analyze.rs:66-149, btf/mod.rs:33-56,
// amd/mod.rs:18-28, kernel.rs:187-262.  The code is written from the published algorithms:
//   * maximum transversal: depth-first augmenting paths with a cheap-assignment pass (Duff 1981);
//   * strong components: Tarjan (1972), non-recursive;
//   * approximate minimum degree: Amestoy, Davis & Duff, SIAM J. Matrix Anal. Appl. 17(4), 1996,
//     and ACM TOMS Algorithm 837 (quotient graph, approximate external degrees, element and
//     aggressive absorption, hash-based supervariable detection, mass elimination, dense-row
//     removal at 10*sqrt(n), postordering with the largest child last);
//   * left-looking LU: Gilbert & Peierls (1988), depth-first reach for each column's pattern.
// The AMD steps are kept in the published order on purpose: the elimination order it returns
// depends on tie-breaking inside those steps, and pivot-order parity with the reference needs
// the same order.
#include "gx_symbolic.hpp"

#include <algorithm>
#include <climits>
#include <cmath>

namespace gx {
namespace {

constexpr int kEmpty = -1;
inline int flip(int x) { return -x - 2; }

}  // namespace

int max_transversal(const CscPattern& a, std::vector<int>& match) {
  const int n = a.n;
  match.assign(n, -1);
  std::vector<int> cheap(a.ap.begin(), a.ap.end() - 1);  // next row to try in the cheap pass
  std::vector<int> visited(n, -1);                        // column -> last search that saw it
  std::vector<int> cstack, pstack;
  int found = 0;
  for (int k = 0; k < n; ++k) {
    // depth-first search for an augmenting path starting at column k
    cstack.assign(1, k);
    pstack.assign(1, a.ap[k]);
    visited[k] = k;
    bool augmented = false;
    while (!cstack.empty() && !augmented) {
      int j = cstack.back();
      // cheap assignment: an unmatched row in column j
      int& c = cheap[j];
      while (c < a.ap[j + 1] && match[a.ai[c]] != -1) ++c;
      if (c < a.ap[j + 1]) {
        int i = a.ai[c++];
        // augment along the stack: each column on the stack takes the row that led onward
        match[i] = j;
        for (int t = (int)cstack.size() - 2; t >= 0; --t) {
          int jj = cstack[t];
          int row = a.ai[pstack[t] - 1];
          match[row] = jj;
        }
        augmented = true;
        break;
      }
      // continue depth first through matched rows
      int& p = pstack.back();
      bool pushed = false;
      while (p < a.ap[j + 1]) {
        int i = a.ai[p++];
        int jn = match[i];
        if (jn >= 0 && visited[jn] != k) {
          visited[jn] = k;
          cstack.push_back(jn);
          pstack.push_back(a.ap[jn]);
          pushed = true;
          break;
        }
      }
      if (!pushed) {
        cstack.pop_back();
        pstack.pop_back();
      }
    }
    if (augmented) ++found;
  }
  return found;
}

int strong_components(const CscPattern& a, const std::vector<int>& colperm, std::vector<int>& p,
                      std::vector<int>& r) {
  const int n = a.n;
  std::vector<int> index(n, -1), low(n, 0), block(n, -1), pos(n, 0);
  std::vector<int> cstack, dstack;
  cstack.reserve(n);
  int stamp = 0, nblocks = 0;
  for (int root = 0; root < n; ++root) {
    if (index[root] >= 0) continue;
    dstack.assign(1, root);
    while (!dstack.empty()) {
      int j = dstack.back();
      int col = colperm[j];
      if (index[j] < 0) {
        index[j] = low[j] = stamp++;
        cstack.push_back(j);
        pos[j] = a.ap[col];
      }
      bool descended = false;
      while (pos[j] < a.ap[col + 1]) {
        int i = a.ai[pos[j]];
        if (index[i] < 0) {
          dstack.push_back(i);
          descended = true;
          break;
        }
        if (block[i] < 0) low[j] = std::min(low[j], index[i]);
        ++pos[j];
      }
      if (descended) continue;
      // j is finished
      dstack.pop_back();
      if (low[j] == index[j]) {
        int i;
        do {
          i = cstack.back();
          cstack.pop_back();
          block[i] = nblocks;
        } while (i != j);
        ++nblocks;
      }
      if (!dstack.empty()) {
        int parent = dstack.back();
        low[parent] = std::min(low[parent], low[j]);
        ++pos[parent];
      }
    }
  }
  r.assign(nblocks + 1, 0);
  for (int i = 0; i < n; ++i) r[block[i] + 1]++;
  for (int b = 0; b < nblocks; ++b) r[b + 1] += r[b];
  std::vector<int> next(r.begin(), r.end() - 1);
  p.assign(n, 0);
  for (int i = 0; i < n; ++i) p[next[block[i]]++] = i;
  return nblocks;
}

namespace {

// Visits every off-diagonal entry pair {i, j} of A + A' exactly once, in the order that the
// adjacency lists are built: column by column, each strictly-upper entry A(j,k) followed by the
// not-yet-seen strictly-lower entries of column j above row k; leftovers last.
template <class F>
void aat_pairs(const CscPattern& a, F&& emit) {
  const int n = a.n;
  std::vector<int> tp(n, 0);
  for (int k = 0; k < n; ++k) {
    int p = a.ap[k];
    const int p2 = a.ap[k + 1];
    while (p < p2) {
      int j = a.ai[p];
      if (j < k) {
        emit(j, k);
        ++p;
      } else if (j == k) {
        ++p;
        break;
      } else {
        break;
      }
      int pj = tp[j];
      const int pj2 = a.ap[j + 1];
      while (pj < pj2) {
        int i = a.ai[pj];
        if (i < k) {
          emit(i, j);
          ++pj;
        } else if (i == k) {
          ++pj;
          break;
        } else {
          break;
        }
      }
      tp[j] = pj;
    }
    tp[k] = p;
  }
  for (int j = 0; j < n; ++j)
    for (int pj = tp[j]; pj < a.ap[j + 1]; ++pj) emit(a.ai[pj], j);
}

int clear_marks(int wflg, int wbig, std::vector<int>& w) {
  if (wflg < 2 || wflg >= wbig) {
    for (auto& x : w)
      if (x != 0) x = 1;
    wflg = 2;
  }
  return wflg;
}

// Postorder of the assembly tree; children of each node in increasing order except that the
// child with the largest front is moved last.  order[e] = position of element e.
void tree_postorder(int n, const std::vector<int>& parent, const std::vector<int>& nv, const std::vector<int>& fsize,
                    std::vector<int>& order) {
  std::vector<int> child(n, kEmpty), sibling(n, kEmpty), stack(n);
  for (int j = n - 1; j >= 0; --j) {
    if (nv[j] > 0 && parent[j] != kEmpty) {
      sibling[j] = child[parent[j]];
      child[parent[j]] = j;
    }
  }
  for (int i = 0; i < n; ++i) {
    if (nv[i] > 0 && child[i] != kEmpty) {
      int fprev = kEmpty, maxf = kEmpty, bigprev = kEmpty, big = kEmpty;
      for (int f = child[i]; f != kEmpty; f = sibling[f]) {
        if (fsize[f] >= maxf) {
          maxf = fsize[f];
          bigprev = fprev;
          big = f;
        }
        fprev = f;
      }
      int fnext = sibling[big];
      if (fnext != kEmpty) {
        if (bigprev == kEmpty) child[i] = fnext;
        else sibling[bigprev] = fnext;
        sibling[big] = kEmpty;
        sibling[fprev] = big;
      }
    }
  }
  order.assign(n, kEmpty);
  int k = 0;
  for (int root = 0; root < n; ++root) {
    if (parent[root] != kEmpty || nv[root] <= 0) continue;
    int head = 0;
    stack[0] = root;
    while (head >= 0) {
      int i = stack[head];
      if (child[i] != kEmpty) {
        for (int f = child[i]; f != kEmpty; f = sibling[f]) ++head;
        int h = head;
        for (int f = child[i]; f != kEmpty; f = sibling[f]) stack[h--] = f;
        child[i] = kEmpty;
      } else {
        --head;
        order[i] = k++;
      }
    }
  }
}

}  // namespace

std::vector<int> amd_order(const CscPattern& a) {
  const int n = a.n;
  std::vector<int> perm(n);
  if (n == 0) return perm;
  // degrees of A + A' and the quotient-graph adjacency lists
  std::vector<int> len(n, 0);
  aat_pairs(a, [&](int i, int j) { ++len[i]; ++len[j]; });
  long long nzaat = 0;
  for (int v : len) nzaat += v;
  const int iwlen = (int)(nzaat + nzaat / 5 + n);
  std::vector<int> pe(n), iw(std::max(iwlen, 1)), sp(n);
  {
    int s = 0;
    for (int i = 0; i < n; ++i) {
      pe[i] = sp[i] = s;
      s += len[i];
    }
  }
  aat_pairs(a, [&](int i, int j) {
    iw[sp[i]++] = j;
    iw[sp[j]++] = i;
  });
  int pfree = (int)nzaat;

  std::vector<int> nv(n, 1), next(n, kEmpty), last(n, kEmpty), head(n, kEmpty), elen(n, 0), degree(len), w(n, 1);
  int dense = (int)(10.0 * std::sqrt((double)n));
  dense = std::max(16, dense);
  dense = std::min(n, dense);
  const int wbig = INT_MAX - n;
  int wflg = clear_marks(0, wbig, w);
  int mindeg = 0, nel = 0, lemax = 0;

  for (int i = 0; i < n; ++i) {
    int deg = degree[i];
    if (deg == 0) {
      elen[i] = flip(1);
      ++nel;
      pe[i] = kEmpty;
      w[i] = 0;
    } else if (deg > dense) {
      nv[i] = 0;
      elen[i] = kEmpty;
      ++nel;
      pe[i] = kEmpty;
    } else {
      int inext = head[deg];
      if (inext != kEmpty) last[inext] = i;
      next[i] = inext;
      head[deg] = i;
    }
  }

  auto unlink = [&](int i) {
    int ilast = last[i], inext = next[i];
    if (inext != kEmpty) last[inext] = ilast;
    if (ilast != kEmpty) next[ilast] = inext;
    else head[degree[i]] = inext;
  };

  while (nel < n) {
    // pivot of minimum approximate degree
    int deg = mindeg;
    int me = kEmpty;
    for (; deg < n; ++deg) {
      me = head[deg];
      if (me != kEmpty) break;
    }
    mindeg = deg;
    {
      int inext = next[me];
      if (inext != kEmpty) last[inext] = kEmpty;
      head[deg] = inext;
    }
    const int elenme = elen[me];
    int nvpiv = nv[me];
    nel += nvpiv;

    // form the new element Lme
    nv[me] = -nvpiv;
    int degme = 0;
    int pme1, pme2;
    if (elenme == 0) {
      pme1 = pe[me];
      pme2 = pme1 - 1;
      const int pend = pme1 + len[me];
      for (int p = pme1; p < pend; ++p) {
        int i = iw[p];
        int nvi = nv[i];
        if (nvi > 0) {
          degme += nvi;
          nv[i] = -nvi;
          iw[++pme2] = i;
          unlink(i);
        }
      }
    } else {
      int p = pe[me];
      pme1 = pfree;
      const int slenme = len[me] - elenme;
      for (int knt1 = 1; knt1 <= elenme + 1; ++knt1) {
        int e, pj, ln;
        if (knt1 > elenme) {
          e = me;
          pj = p;
          ln = slenme;
        } else {
          e = iw[p++];
          pj = pe[e];
          ln = len[e];
        }
        for (int knt2 = 1; knt2 <= ln; ++knt2) {
          int i = iw[pj++];
          int nvi = nv[i];
          if (nvi <= 0) continue;
          if (pfree >= iwlen) {
            // out of room: compact every live list to the front, keeping each list's order
            pe[me] = p;
            len[me] -= knt1;
            if (len[me] == 0) pe[me] = kEmpty;
            pe[e] = pj;
            len[e] = ln - knt2;
            if (len[e] == 0) pe[e] = kEmpty;
            for (int j = 0; j < n; ++j) {
              int pn = pe[j];
              if (pn >= 0) {
                pe[j] = iw[pn];
                iw[pn] = flip(j);
              }
            }
            int psrc = 0, pdst = 0;
            const int pend = pme1 - 1;
            while (psrc <= pend) {
              int j = flip(iw[psrc++]);
              if (j >= 0) {
                iw[pdst] = pe[j];
                pe[j] = pdst++;
                for (int t = 0; t <= len[j] - 2; ++t) iw[pdst++] = iw[psrc++];
              }
            }
            int p1 = pdst;
            for (psrc = pme1; psrc <= pfree - 1; ++psrc) iw[pdst++] = iw[psrc];
            pme1 = p1;
            pfree = pdst;
            pj = pe[e];
            p = pe[me];
          }
          degme += nvi;
          nv[i] = -nvi;
          iw[pfree++] = i;
          unlink(i);
        }
        if (e != me) {
          pe[e] = flip(me);  // e is absorbed into me
          w[e] = 0;
        }
      }
      pme2 = pfree - 1;
    }
    degree[me] = degme;
    pe[me] = pme1;
    len[me] = pme2 - pme1 + 1;
    elen[me] = flip(nvpiv + degme);
    wflg = clear_marks(wflg, wbig, w);

    // |Le \ Lme| for every element e adjacent to Lme
    for (int pme = pme1; pme <= pme2; ++pme) {
      int i = iw[pme];
      int eln = elen[i];
      if (eln <= 0) continue;
      int nvi = -nv[i];
      int wnvi = wflg - nvi;
      for (int p = pe[i]; p <= pe[i] + eln - 1; ++p) {
        int e = iw[p];
        int we = w[e];
        if (we >= wflg) we -= nvi;
        else if (we != 0) we = degree[e] + wnvi;
        w[e] = we;
      }
    }

    // approximate degrees, element absorption, mass elimination, hashing
    for (int pme = pme1; pme <= pme2; ++pme) {
      int i = iw[pme];
      int p1 = pe[i];
      int p2 = p1 + elen[i] - 1;
      int pn = p1;
      unsigned int hash = 0;
      int d = 0;
      for (int p = p1; p <= p2; ++p) {
        int e = iw[p];
        int we = w[e];
        if (we == 0) continue;
        int dext = we - wflg;
        if (dext > 0) {
          d += dext;
          iw[pn++] = e;
          hash += (unsigned int)e;
        } else {
          pe[e] = flip(me);  // aggressive absorption: Le is a subset of Lme
          w[e] = 0;
        }
      }
      elen[i] = pn - p1 + 1;
      int p3 = pn;
      int p4 = p1 + len[i];
      for (int p = p2 + 1; p < p4; ++p) {
        int j = iw[p];
        int nvj = nv[j];
        if (nvj > 0) {
          d += nvj;
          iw[pn++] = j;
          hash += (unsigned int)j;
        }
      }
      if (elen[i] == 1 && p3 == pn) {
        // i is adjacent only to me: eliminate it together with me
        pe[i] = flip(me);
        int nvi = -nv[i];
        degme -= nvi;
        nvpiv += nvi;
        nel += nvi;
        nv[i] = 0;
        elen[i] = kEmpty;
      } else {
        degree[i] = std::min(degree[i], d);
        iw[pn] = iw[p3];
        iw[p3] = iw[p1];
        iw[p1] = me;
        len[i] = pn - p1 + 1;
        hash %= (unsigned int)n;
        int j = head[hash];
        if (j <= kEmpty) {
          next[i] = flip(j);
          head[hash] = flip(i);
        } else {
          next[i] = last[j];
          last[j] = i;
        }
        last[i] = (int)hash;
      }
    }
    degree[me] = degme;
    lemax = std::max(lemax, degme);
    wflg += lemax;
    wflg = clear_marks(wflg, wbig, w);

    // supervariable detection: variables with identical adjacency merge into one
    for (int pme = pme1; pme <= pme2; ++pme) {
      int i = iw[pme];
      if (nv[i] >= 0) continue;
      unsigned int hash = (unsigned int)last[i];
      int j = head[hash];
      if (j == kEmpty) {
        i = kEmpty;
      } else if (j < kEmpty) {
        i = flip(j);
        head[hash] = kEmpty;
      } else {
        i = last[j];
        last[j] = kEmpty;
      }
      while (i != kEmpty && next[i] != kEmpty) {
        int ln = len[i];
        int eln = elen[i];
        for (int p = pe[i] + 1; p <= pe[i] + ln - 1; ++p) w[iw[p]] = wflg;
        int jlast = i;
        j = next[i];
        while (j != kEmpty) {
          bool same = (len[j] == ln) && (elen[j] == eln);
          for (int p = pe[j] + 1; same && p <= pe[j] + ln - 1; ++p)
            if (w[iw[p]] != wflg) same = false;
          if (same) {
            pe[j] = flip(i);
            nv[i] += nv[j];
            nv[j] = 0;
            elen[j] = kEmpty;
            j = next[j];
            next[jlast] = j;
          } else {
            jlast = j;
            j = next[j];
          }
        }
        ++wflg;
        i = next[i];
      }
    }

    // back into the degree lists; drop non-principal variables from Lme
    int p = pme1;
    const int nleft = n - nel;
    for (int pme = pme1; pme <= pme2; ++pme) {
      int i = iw[pme];
      int nvi = -nv[i];
      if (nvi <= 0) continue;
      nv[i] = nvi;
      int d = std::min(degree[i] + degme - nvi, nleft - nvi);
      int inext = head[d];
      if (inext != kEmpty) last[inext] = i;
      next[i] = inext;
      last[i] = kEmpty;
      head[d] = i;
      mindeg = std::min(mindeg, d);
      degree[i] = d;
      iw[p++] = i;
    }
    nv[me] = nvpiv;
    len[me] = p - pme1;
    if (len[me] == 0) {
      pe[me] = kEmpty;
      w[me] = 0;
    }
    if (elenme != 0) pfree = p;
  }

  // assembly tree: parents of elements, principal variable of every absorbed variable
  for (int i = 0; i < n; ++i) pe[i] = flip(pe[i]);
  for (int i = 0; i < n; ++i) elen[i] = flip(elen[i]);
  for (int i = 0; i < n; ++i) {
    if (nv[i] != 0) continue;
    int j = pe[i];
    if (j == kEmpty) continue;  // dense row: ordered last
    while (nv[j] == 0) j = pe[j];
    int e = j;
    j = i;
    while (nv[j] == 0) {
      int jn = pe[j];
      pe[j] = e;
      j = jn;
    }
  }
  std::vector<int> order;
  tree_postorder(n, pe, nv, elen, order);
  std::fill(head.begin(), head.end(), kEmpty);
  std::fill(next.begin(), next.end(), kEmpty);
  for (int e = 0; e < n; ++e)
    if (order[e] != kEmpty) head[order[e]] = e;
  int k = 0;
  for (int t = 0; t < n; ++t) {
    int e = head[t];
    if (e == kEmpty) break;
    next[e] = k;
    k += nv[e];
  }
  for (int i = 0; i < n; ++i) {
    if (nv[i] != 0) continue;
    int e = pe[i];
    if (e != kEmpty) next[i] = next[e]++;
    else next[i] = k++;
  }
  for (int i = 0; i < n; ++i) perm[next[i]] = i;
  return perm;
}

bool analyze(const CscPattern& a, Ordering& ord) {
  const int n = a.n;
  std::vector<int> match;
  int nm = max_transversal(a, match);
  if (nm < n) return false;  // structurally singular: the Newton step cannot be factored
  std::vector<int> qm(n);
  for (int i = 0; i < n; ++i) qm[i] = match[i];  // node i (row i on the diagonal) <- column match[i]
  std::vector<int> pb, r;
  int nblocks = strong_components(a, qm, pb, r);
  std::vector<int> qb(n), pinvb(n);
  for (int k = 0; k < n; ++k) {
    qb[k] = qm[pb[k]];
    pinvb[pb[k]] = k;
  }
  ord.p.assign(n, 0);
  ord.q.assign(n, 0);
  ord.r = r;
  for (int b = 0; b < nblocks; ++b) {
    const int k1 = r[b], k2 = r[b + 1], nk = k2 - k1;
    std::vector<int> pblk(nk);
    if (nk <= 3) {
      for (int k = 0; k < nk; ++k) pblk[k] = k;
    } else {
      CscPattern c;
      c.n = nk;
      c.ap.assign(nk + 1, 0);
      for (int k = 0; k < nk; ++k) {
        int col = qb[k + k1];
        for (int p = a.ap[col]; p < a.ap[col + 1]; ++p) {
          int nr = pinvb[a.ai[p]];
          if (nr >= k1 && nr < k2) c.ai.push_back(nr - k1);
        }
        std::sort(c.ai.begin() + c.ap[k], c.ai.end());
        c.ap[k + 1] = (int)c.ai.size();
      }
      pblk = amd_order(c);
    }
    for (int k = 0; k < nk; ++k) {
      ord.q[k + k1] = qb[pblk[k] + k1];
      ord.p[k + k1] = pb[pblk[k] + k1];
    }
  }
  return true;
}

namespace {

// Packs per-column U (ascending rows < k) / diagonal / L (ascending rows > k) lists.
void pack(int n, const std::vector<std::vector<int>>& ucols, const std::vector<std::vector<int>>& lcols, LuPattern& lu) {
  lu.n = n;
  lu.cp.assign(n + 1, 0);
  lu.ulen.assign(n, 0);
  lu.row.clear();
  for (int k = 0; k < n; ++k) {
    std::vector<int> u = ucols[k], l = lcols[k];
    std::sort(u.begin(), u.end());
    std::sort(l.begin(), l.end());
    lu.ulen[k] = (int)u.size();
    lu.row.insert(lu.row.end(), u.begin(), u.end());
    lu.row.push_back(k);
    lu.row.insert(lu.row.end(), l.begin(), l.end());
    lu.cp[k + 1] = (int)lu.row.size();
  }
}

}  // namespace

bool factor_pivot(const CscPattern& a, const double* ax, const Ordering& ord, double tol, LuPattern& lu) {
  const int n = a.n;
  std::vector<int> pinv(n, -1), mark(n, -1), stack, pstack, reach;
  std::vector<double> x(n, 0.0);
  // L columns with original row indices and values while factoring
  std::vector<std::vector<int>> lrow(n), ucol(n);
  std::vector<std::vector<double>> lval(n);
  lu.q = ord.q;
  lu.pnum.assign(n, -1);
  // Preferred ("diagonal") row of every step, and the step that currently prefers each row.
  // When a step pivots on another step's preferred row, the displaced row becomes that other
  // step's preference (KLU's diagonal bookkeeping, src/klu_native/kernel.rs:388-395).
  std::vector<int> pref(ord.p), pref_of(n, -1);
  for (int k = 0; k < n; ++k) pref_of[pref[k]] = k;
  for (int k = 0; k < n; ++k) {
    const int col = ord.q[k];
    reach.clear();
    for (int p = a.ap[col]; p < a.ap[col + 1]; ++p) {
      int r0 = a.ai[p];
      if (mark[r0] == k) continue;
      mark[r0] = k;
      stack.assign(1, r0);
      pstack.assign(1, 0);
      while (!stack.empty()) {
        int v = stack.back();
        int j = pinv[v];
        if (j < 0) {
          reach.push_back(v);
          stack.pop_back();
          pstack.pop_back();
          continue;
        }
        int& t = pstack.back();
        bool pushed = false;
        while (t < (int)lrow[j].size()) {
          int w = lrow[j][t++];
          if (mark[w] != k) {
            mark[w] = k;
            stack.push_back(w);
            pstack.push_back(0);
            pushed = true;
            break;
          }
        }
        if (!pushed) {
          reach.push_back(v);
          stack.pop_back();
          pstack.pop_back();
        }
      }
    }
    for (int p = a.ap[col]; p < a.ap[col + 1]; ++p) x[a.ai[p]] = ax[p];
    for (int t = (int)reach.size() - 1; t >= 0; --t) {
      int v = reach[t];
      int j = pinv[v];
      if (j < 0) continue;
      double xj = x[v];
      for (size_t e = 0; e < lrow[j].size(); ++e) x[lrow[j][e]] -= lval[j][e] * xj;
      ucol[k].push_back(j);
      x[v] = 0.0;
    }
    // pivot: largest candidate, unless the preferred row's candidate passes the tolerance
    int piv = -1, diag = -1;
    double amax = -1.0;
    for (int t = (int)reach.size() - 1; t >= 0; --t) {
      int v = reach[t];
      if (pinv[v] >= 0) continue;
      double av = std::fabs(x[v]);
      if (av > amax) {
        amax = av;
        piv = v;
      }
      if (v == pref[k]) diag = v;
    }
    if (piv < 0) return false;
    if (diag >= 0 && std::fabs(x[diag]) >= tol * amax) piv = diag;
    double pv = x[piv];
    if (pv == 0.0 || !std::isfinite(pv)) return false;
    const int diagrow = pref[k];
    if (piv != diagrow && pinv[diagrow] < 0) {
      const int kbar = pref_of[piv];
      pref[kbar] = diagrow;
      pref_of[diagrow] = kbar;
    }
    pref[k] = piv;
    pref_of[piv] = k;
    pinv[piv] = k;
    lu.pnum[k] = piv;
    x[piv] = 0.0;
    for (int t = (int)reach.size() - 1; t >= 0; --t) {
      int v = reach[t];
      if (pinv[v] >= 0) continue;
      lrow[k].push_back(v);
      lval[k].push_back(x[v] / pv);
      x[v] = 0.0;
    }
  }
  lu.pinv = pinv;
  std::vector<std::vector<int>> lcols(n);
  for (int k = 0; k < n; ++k) {
    lcols[k].reserve(lrow[k].size());
    for (int v : lrow[k]) lcols[k].push_back(pinv[v]);
  }
  pack(n, ucol, lcols, lu);
  return true;
}

bool pattern_from_order(const CscPattern& a, const std::vector<int>& q, const std::vector<int>& pnum, LuPattern& lu) {
  const int n = a.n;
  if ((int)q.size() != n || (int)pnum.size() != n) return false;
  std::vector<int> pinv(n, -1);
  for (int k = 0; k < n; ++k) {
    if (pnum[k] < 0 || pnum[k] >= n || pinv[pnum[k]] >= 0) return false;
    pinv[pnum[k]] = k;
  }
  std::vector<int> qseen(n, 0);
  for (int k = 0; k < n; ++k) {
    if (q[k] < 0 || q[k] >= n || qseen[q[k]]) return false;
    qseen[q[k]] = 1;
  }
  // symbolic left-looking factorization of A(pnum, q) in pivot-step space
  std::vector<std::vector<int>> lcols(n), ucols(n);
  std::vector<int> mark(n, -1), stack, pstack, reach;
  for (int k = 0; k < n; ++k) {
    reach.clear();
    bool has_diag = false;
    for (int p = a.ap[q[k]]; p < a.ap[q[k] + 1]; ++p) {
      int r0 = pinv[a.ai[p]];
      if (mark[r0] == k) continue;
      mark[r0] = k;
      stack.assign(1, r0);
      pstack.assign(1, 0);
      while (!stack.empty()) {
        int v = stack.back();
        if (v >= k) {
          reach.push_back(v);
          stack.pop_back();
          pstack.pop_back();
          continue;
        }
        int& t = pstack.back();
        bool pushed = false;
        while (t < (int)lcols[v].size()) {
          int w = lcols[v][t++];
          if (mark[w] != k) {
            mark[w] = k;
            stack.push_back(w);
            pstack.push_back(0);
            pushed = true;
            break;
          }
        }
        if (!pushed) {
          reach.push_back(v);
          stack.pop_back();
          pstack.pop_back();
        }
      }
    }
    for (int v : reach) {
      if (v < k) ucols[k].push_back(v);
      else if (v > k) lcols[k].push_back(v);
      else has_diag = true;
    }
    if (!has_diag) return false;  // structurally zero pivot
    std::sort(lcols[k].begin(), lcols[k].end());
  }
  lu.q = q;
  lu.pnum = pnum;
  lu.pinv = pinv;
  pack(n, ucols, lcols, lu);
  return true;
}

}  // namespace gx
