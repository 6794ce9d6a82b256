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
//     removal at 10*sqrt(n), postordering with the largest child last), on per-node lists of this
//     file's own (no shared workspace, no compaction);
//   * left-looking LU: Gilbert & Peierls (1988), depth-first reach for each column's pattern.
// The elimination order depends on tie-breaking (bucket and list orders), and pivot-order parity
// with the reference needs the same order, so those orders follow the published method.
#include "gx_symbolic.hpp"

#include <algorithm>
#include <climits>
#include <cmath>

namespace gx {
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

// ---------------------------------------------------------------- A + A'
// Neighbour lists of the pattern of A + A' (diagonal dropped, every pair once at each end), in the
// order in which the published AMD code builds them: columns in increasing order; in column k each
// strictly upper entry A(j, k), by increasing j, adds the pair {j, k} and then releases the strictly
// lower entries A(i, j) of column j with i < k that are still pending (A(k, j), the mirror of
// A(j, k), is passed over); lower entries no column releases come last, column by column.  The
// minimum-degree order breaks ties by list position, so this order is part of its result.
std::vector<std::vector<int>> symmetric_adjacency(const CscPattern& a) {
  const int n = a.n;
  std::vector<std::vector<int>> adj(n);
  auto link = [&adj](int u, int v) {
    adj[u].push_back(v);
    adj[v].push_back(u);
  };
  std::vector<int> pending(n, 0);  // column j: its first strictly lower entry not released yet
  for (int k = 0; k < n; ++k) {
    const int end = a.ap[k + 1];
    int at = a.ap[k];
    for (; at < end && a.ai[at] < k; ++at) {
      const int j = a.ai[at];
      link(j, k);
      int& c = pending[j];
      const int cend = a.ap[j + 1];
      for (; c < cend && a.ai[c] < k; ++c) link(a.ai[c], j);
      if (c < cend && a.ai[c] == k) ++c;
    }
    if (at < end && a.ai[at] == k) ++at;
    pending[k] = at;
  }
  for (int j = 0; j < n; ++j)
    for (int c = pending[j]; c < a.ap[j + 1]; ++c) link(a.ai[c], j);
  return adj;
}

// ---------------------------------------------------------------- degree buckets
// Variables grouped by approximate degree, each group a doubly linked list: a variable enters at the
// front, and the next pivot is the front of the lowest non-empty group.
class DegreeBuckets {
 public:
  explicit DegreeBuckets(int n) : front_(n + 1, -1), before_(n, -1), after_(n, -1), group_(n, -1) {}
  void add(int v, int d) {
    const int f = front_[d];
    before_[v] = -1;
    after_[v] = f;
    if (f >= 0) before_[f] = v;
    front_[d] = v;
    group_[v] = d;
  }
  void remove(int v) {
    const int b = before_[v], a = after_[v];
    if (a >= 0) before_[a] = b;
    if (b >= 0) after_[b] = a;
    else front_[group_[v]] = a;
    group_[v] = -1;
  }
  int front(int d) const { return front_[d]; }

 private:
  std::vector<int> front_, before_, after_, group_;
};

// ---------------------------------------------------------------- assembly-tree postorder
// Position of every element in a postorder of the assembly tree: roots by increasing index, the
// children of a node by increasing index except that the child with the largest front (the last
// such child on ties) is visited last.
std::vector<int> assembly_postorder(const std::vector<int>& parent, const std::vector<char>& is_node,
                                    const std::vector<int>& front) {
  const int n = (int)parent.size();
  std::vector<std::vector<int>> kids(n);
  for (int v = 0; v < n; ++v)
    if (is_node[v] && parent[v] >= 0) kids[parent[v]].push_back(v);
  for (auto& c : kids) {
    if (c.size() < 2) continue;
    size_t big = 0;
    for (size_t t = 1; t < c.size(); ++t)
      if (front[c[t]] >= front[c[big]]) big = t;
    const int b = c[big];
    c.erase(c.begin() + (long)big);
    c.push_back(b);
  }
  std::vector<int> pos(n, -1);
  std::vector<std::pair<int, size_t>> walk;  // (node, next child to visit)
  int count = 0;
  for (int root = 0; root < n; ++root) {
    if (!is_node[root] || parent[root] >= 0) continue;
    walk.assign(1, {root, 0});
    while (!walk.empty()) {
      auto& [v, next] = walk.back();
      if (next < kids[v].size()) {
        const int c = kids[v][next++];
        walk.push_back({c, 0});
      } else {
        pos[v] = count++;
        walk.pop_back();
      }
    }
  }
  return pos;
}

}  // namespace

// Approximate minimum degree (Amestoy, Davis & Duff 1996) on the quotient graph of A + A'.
// Every variable keeps its adjacent elements and its adjacent variables as two ordered lists and
// every element its member variables; eliminating a pivot forms a new element from the members of
// the pivot's elements and its own variables, absorbs those elements, and then, for each member:
// approximate external degree |Le \ Lme| of its other elements (aggressive absorption of those with
// none left), mass elimination when the new element is its only neighbour, and a hash of its
// neighbourhood; members with equal neighbourhoods merge into one supervariable.  Rows denser than
// max(16, 10 sqrt n) are left out and ordered last.  Ties follow the list and bucket orders of the
// published method (new element first in each member's element list, that list's former first entry
// and the first variable moved to the ends; buckets filled at the front), which fix the result.
std::vector<int> amd_order(const CscPattern& a) {
  const int n = a.n;
  std::vector<int> perm(n);
  if (n == 0) return perm;
  std::vector<std::vector<int>> vars = symmetric_adjacency(a);  // variables adjacent to a variable
  std::vector<std::vector<int>> elems(n);                       // elements adjacent to a variable
  std::vector<std::vector<int>> members(n);                     // variables of an element
  std::vector<int> weight(n, 1);      // supervariable size; 0: merged into another or dense
  std::vector<int> degree(n, 0);      // variables: approximate degree; elements: external degree
  std::vector<int> parent(n, -1);     // absorbing element / supervariable
  std::vector<int> front(n, 0);       // elements: front size (pivots + external degree)
  std::vector<char> alive(n, 0), in_front(n, 0);
  std::vector<int> ext(n, 0), ext_stamp(n, 0), mark(n, 0), hash_of(n, 0);
  std::vector<std::vector<int>> hash_bucket(n);
  int stamp = 0, mark_stamp = 0;
  DegreeBuckets buckets(n);

  const int dense = std::min(n, std::max(16, (int)(10.0 * std::sqrt((double)n))));
  int eliminated = 0;
  for (int v = 0; v < n; ++v) {
    const int d = (int)vars[v].size();
    degree[v] = d;
    if (d == 0) {  // no neighbour: an element of its own at once (a root)
      front[v] = 1;
      ++eliminated;
    } else if (d > dense) {  // dense: left out, ordered last
      weight[v] = 0;
      ++eliminated;
    } else {
      buckets.add(v, d);
    }
  }

  std::vector<int> lme;
  int mindeg = 0;
  while (eliminated < n) {
    // the pivot: a variable of least approximate degree
    while (buckets.front(mindeg) < 0) ++mindeg;
    const int me = buckets.front(mindeg);
    buckets.remove(me);
    int npiv = weight[me];
    eliminated += npiv;
    in_front[me] = 1;

    // the new element: the members of the pivot's elements (each absorbed into it), then the
    // pivot's own variables
    lme.clear();
    long long new_elem_degree = 0;
    auto take = [&](int v) {
      if (weight[v] > 0 && !in_front[v]) {
        new_elem_degree += weight[v];
        in_front[v] = 1;
        lme.push_back(v);
        buckets.remove(v);
      }
    };
    for (int e : elems[me]) {
      for (int v : members[e]) take(v);
      parent[e] = me;
      alive[e] = 0;
      std::vector<int>().swap(members[e]);
    }
    for (int v : vars[me]) take(v);
    std::vector<int>().swap(elems[me]);
    std::vector<int>().swap(vars[me]);

    // |Le \ Lme| of every element adjacent to the new element's members
    ++stamp;
    for (int v : lme)
      for (int e : elems[v]) {
        if (!alive[e]) continue;
        if (ext_stamp[e] != stamp) {
          ext_stamp[e] = stamp;
          ext[e] = degree[e];
        }
        ext[e] -= weight[v];
      }

    // degrees, absorption, mass elimination and neighbourhood hashes of the members
    for (int v : lme) {
      long long deg = 0;
      unsigned long long h = 0;
      std::vector<int>& el = elems[v];
      size_t ke = 0;
      for (int e : el) {
        if (!alive[e]) continue;
        if (ext[e] > 0) {
          deg += ext[e];
          h += (unsigned long long)e;
          el[ke++] = e;
        } else {  // no variable of e is outside the new element: absorbed into it
          parent[e] = me;
          alive[e] = 0;
          std::vector<int>().swap(members[e]);
        }
      }
      el.resize(ke);
      std::vector<int>& vl = vars[v];
      size_t kv = 0;
      for (int u : vl)
        if (weight[u] > 0 && !in_front[u]) {
          deg += weight[u];
          h += (unsigned long long)u;
          vl[kv++] = u;
        }
      vl.resize(kv);
      if (ke == 0 && kv == 0) {  // only the new element is left: eliminated with the pivot
        parent[v] = me;
        new_elem_degree -= weight[v];
        npiv += weight[v];
        eliminated += weight[v];
        weight[v] = 0;
        std::vector<int>().swap(el);
        std::vector<int>().swap(vl);
        continue;
      }
      degree[v] = (int)std::min<long long>(degree[v], deg);
      if (ke > 0) {
        el.push_back(el.front());
        el.front() = me;
      } else {
        el.push_back(me);
      }
      if (kv > 1) std::rotate(vl.begin(), vl.begin() + 1, vl.end());
      hash_of[v] = (int)(h % (unsigned long long)n);
      hash_bucket[hash_of[v]].push_back(v);
    }
    degree[me] = (int)new_elem_degree;

    // supervariables: members with the same elements and variables (besides the new element) merge
    // into the member hashed last among them; buckets are searched in member order
    std::vector<int> chain;
    for (int v : lme) {
      if (weight[v] == 0) continue;
      std::vector<int>& bucket = hash_bucket[hash_of[v]];
      if (bucket.empty()) continue;
      chain.assign(bucket.rbegin(), bucket.rend());
      bucket.clear();
      for (size_t x = 0; x + 1 < chain.size(); ++x) {
        const int s = chain[x];
        ++mark_stamp;
        for (size_t t = 1; t < elems[s].size(); ++t) mark[elems[s][t]] = mark_stamp;
        for (int u : vars[s]) mark[u] = mark_stamp;
        size_t keep = x + 1;
        for (size_t y = x + 1; y < chain.size(); ++y) {
          const int j = chain[y];
          bool same = elems[j].size() == elems[s].size() && vars[j].size() == vars[s].size();
          for (size_t t = 1; same && t < elems[j].size(); ++t) same = mark[elems[j][t]] == mark_stamp;
          for (size_t t = 0; same && t < vars[j].size(); ++t) same = mark[vars[j][t]] == mark_stamp;
          if (same) {
            parent[j] = s;
            weight[s] += weight[j];
            weight[j] = 0;
            std::vector<int>().swap(elems[j]);
            std::vector<int>().swap(vars[j]);
          } else {
            chain[keep++] = j;
          }
        }
        chain.resize(keep);
      }
    }

    // the surviving members return to the buckets with their degrees including the new element
    const int left = n - eliminated;
    std::vector<int>& mem = members[me];
    for (int v : lme) {
      in_front[v] = 0;
      const int w = weight[v];
      if (w == 0) continue;
      const int d = (int)std::min<long long>((long long)degree[v] + new_elem_degree - w, (long long)left - w);
      buckets.add(v, d);
      mindeg = std::min(mindeg, d);
      degree[v] = d;
      mem.push_back(v);
    }
    in_front[me] = 0;
    weight[me] = npiv;
    front[me] = npiv + (int)new_elem_degree;
    alive[me] = mem.empty() ? 0 : 1;
    if (mem.empty()) parent[me] = -1;
  }

  // every merged variable joins the element its supervariable ended in
  for (int v = 0; v < n; ++v) {
    if (weight[v] != 0 || parent[v] < 0) continue;
    int e = parent[v];
    while (weight[e] == 0) e = parent[e];
    for (int u = v; weight[u] == 0;) {
      const int up = parent[u];
      parent[u] = e;
      u = up;
    }
  }
  std::vector<char> node(n, 0);
  for (int v = 0; v < n; ++v) node[v] = weight[v] > 0;
  const std::vector<int> post = assembly_postorder(parent, node, front);
  // an element's slots: its merged variables in increasing order, then the element; dense rows last
  std::vector<int> by_post(n, -1), slot(n, 0);
  for (int v = 0; v < n; ++v)
    if (node[v]) by_post[post[v]] = v;
  int next = 0;
  for (int k = 0; k < n && by_post[k] >= 0; ++k) {
    slot[by_post[k]] = next;
    next += weight[by_post[k]];
  }
  std::vector<int> where(n, 0);
  for (int v = 0; v < n; ++v) {
    if (node[v]) continue;
    where[v] = parent[v] >= 0 ? slot[parent[v]]++ : next++;
  }
  for (int v = 0; v < n; ++v)
    if (node[v]) where[v] = slot[v];
  for (int v = 0; v < n; ++v) perm[where[v]] = v;
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
