// CPU oracle for the batched Newton-Raphson AC power flow.
//
// TEST INFRASTRUCTURE ONLY.  This file restates the reference algorithm on the CPU so that the
// HIP product can be checked against it.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load the library built from it, and only as the checker / the timed CPU
// baseline.  The product (gridoxide_amd/) never calls into it.
//
// Reference functions restated here (file:line in the reference):
//   Y-bus build          build_ybus             src/network.rs:225-247
//                        branch_calc_param (1,1) src/network.rs:339-360
//                        stamp_transformers     src/network.rs:364-373
//                        YBus::finish           src/network.rs:52-71 (duplicates summed in push order)
//   power_injections     src/network.rs:680-701 (I = Y*V, S_k = V_k*conj(I_k))
//   effective_injection  src/network.rs:600-611 (ZIP terms summed per kind, evaluated at the bus's vm)
//   Jacobian triplets    build_jacobian_triplets src/solver.rs:818-888 (H/N then M/L emission order)
//   NR loop              newton_raphson_cached  src/solver.rs:639-736 (history, status, update)
//   sparse LU            KLU factor with diagonal-preference partial pivoting at tol 1e-3
//                        (src/klu_native/kernel.rs:187-262), in-place refactor with the pivot order
//                        kept (src/klu_native/kernel.rs:476-524) and the permuted solve
//                        (src/klu_native/solve.rs:56-102).  The oracle orders columns bus-major
//                        (angle then magnitude unknown of each bus) instead of BTF+AMD: a different
//                        but valid elimination order, so results agree with the reference to
//                        rounding, not bit-for-bit.
//   linear_initial_guess src/network.rs:627-678 (complex system, solved by the same LU in complex)
//   islands              connected_components / classify / mark_unreferenced_islands
//                        src/network.rs:121-223, resolve_pending src/solver.rs:234-265
//   BatchSolver          src/batch.rs:198-289 (overrides already applied by the caller; pivots fixed
//                        by the first factorization, as a one-worker PersistentSolver does)
//
// Complex arithmetic follows num-complex 0.4.6's formulas (mul: (ac-bd, ad+bc); div by norm_sqr;
// from_polar: (r cos t, r sin t); norm: hypot; arg: atan2), built with -ffp-contract=off.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <thread>
#include <utility>
#include <vector>

namespace {

struct cx {
  double re, im;
};
inline cx cadd(cx a, cx b) { return {a.re + b.re, a.im + b.im}; }
inline cx csub(cx a, cx b) { return {a.re - b.re, a.im - b.im}; }
inline cx cmul(cx a, cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline cx cdiv(cx a, cx b) {
  double ns = b.re * b.re + b.im * b.im;
  return {(a.re * b.re + a.im * b.im) / ns, (a.im * b.re - a.re * b.im) / ns};
}
inline cx cconj(cx a) { return {a.re, -a.im}; }
inline cx cneg(cx a) { return {-a.re, -a.im}; }
inline cx cscale(cx a, double s) { return {a.re * s, a.im * s}; }
inline cx polar(double r, double t) { return {r * std::cos(t), r * std::sin(t)}; }
inline double cabs_(cx a) { return std::hypot(a.re, a.im); }

// ---------------------------------------------------------------- Y-bus
struct YBus {
  int n = 0;
  std::vector<int> rowptr, col;
  std::vector<cx> val;
};

// Scalar helpers so the LU below can be written once for real and complex values.
inline double sabs(double a) { return std::fabs(a); }
inline double sabs(cx a) { return cabs_(a); }
inline bool sisfinite(double a) { return std::isfinite(a); }
inline bool sisfinite(cx a) { return std::isfinite(a.re) && std::isfinite(a.im); }
inline bool siszero(double a) { return a == 0.0; }
inline bool siszero(cx a) { return a.re == 0.0 && a.im == 0.0; }
inline double ssub_mul(double x, double l, double u) { return x - l * u; }
inline cx ssub_mul(cx x, cx l, cx u) { return csub(x, cmul(l, u)); }
inline double sdiv(double a, double b) { return a / b; }
inline cx sdiv(cx a, cx b) { return cdiv(a, b); }
template <class T> inline T szero();
template <> inline double szero<double>() { return 0.0; }
template <> inline cx szero<cx>() { return {0.0, 0.0}; }

// ---------------------------------------------------------------- sparse LU (KLU semantics)
// Left-looking Gilbert-Peierls LU of A(:, q) with row partial pivoting that prefers the
// diagonal candidate when |a_diag| >= tol * max|a| (tol = 1e-3, KLU's default).  The pivot order
// found by factor() is kept by refactor(), as KLU's refactor does.
template <class T>
struct SparseLU {
  int n = 0;
  std::vector<int> q;         // column order: pivot step k factors original column q[k]
  std::vector<int> pinv;      // original row -> pivot step
  std::vector<int> prow;      // pivot step -> original row
  std::vector<int> lp, li;    // L (unit, strictly lower), rows in pivot-step space after finish
  std::vector<T> lx;
  std::vector<int> up, ui;    // U strictly upper, rows = pivot steps, stored in topological order
  std::vector<T> ux;
  std::vector<T> udiag;
  std::vector<T> x;           // dense work vector indexed by original row
  bool ok = false;

  // CSC input: ap[n+1], ai, ax.  diag_row[k] is the preferred (diagonal) row for step k.
  bool factor(int n_, const int* ap, const int* ai, const T* ax, const std::vector<int>& q_,
              const std::vector<int>& diag_row, double tol) {
    n = n_;
    q = q_;
    pinv.assign(n, -1);
    prow.assign(n, -1);
    lp.assign(1, 0);
    up.assign(1, 0);
    li.clear(); lx.clear(); ui.clear(); ux.clear();
    udiag.assign(n, szero<T>());
    x.assign(n, szero<T>());
    std::vector<int> mark(n, -1), stack, out, pstack;
    std::vector<int> lrows_tmp;
    // L columns are kept with original row indices during the factorization (rows are assigned a
    // pivot step only when they are chosen); they are renumbered to pivot steps at the end.
    for (int k = 0; k < n; ++k) {
      int col = q[k];
      // symbolic reach of A(:,col) through the graph of L (columns already factored)
      out.clear();
      for (int p = ap[col]; p < ap[col + 1]; ++p) {
        int r = ai[p];
        if (mark[r] == k) continue;
        // iterative DFS from r
        stack.clear(); pstack.clear();
        stack.push_back(r); pstack.push_back(-1);
        mark[r] = k;
        while (!stack.empty()) {
          int v = stack.back();
          int j = pinv[v];
          int& pp = pstack.back();
          if (j < 0) {  // non-pivotal row: leaf
            out.push_back(v);
            stack.pop_back(); pstack.pop_back();
            continue;
          }
          if (pp < 0) pp = lp[j];
          bool pushed = false;
          while (pp < lp[j + 1]) {
            int w = li[pp++];
            if (mark[w] != k) {
              mark[w] = k;
              stack.push_back(w); pstack.push_back(-1);
              pushed = true;
              break;
            }
          }
          if (!pushed) {
            out.push_back(v);
            stack.pop_back(); pstack.pop_back();
          }
        }
      }
      // scatter A(:,col)
      for (int p = ap[col]; p < ap[col + 1]; ++p) x[ai[p]] = ax[p];
      // numeric triangular solve in topological order (reverse of DFS finish order)
      for (int t = (int)out.size() - 1; t >= 0; --t) {
        int v = out[t];
        int j = pinv[v];
        if (j < 0) continue;
        T xj = x[v];
        for (int p = lp[j]; p < lp[j + 1]; ++p) x[li[p]] = ssub_mul(x[li[p]], lx[p], xj);
      }
      // U(:,k): pivotal rows, in topological order
      for (int t = (int)out.size() - 1; t >= 0; --t) {
        int v = out[t];
        int j = pinv[v];
        if (j < 0) continue;
        ui.push_back(j);
        ux.push_back(x[v]);
        x[v] = szero<T>();
      }
      up.push_back((int)ui.size());
      // pivot among non-pivotal rows, diagonal preference
      int piv = -1;
      double amax = -1.0;
      int dpos = -1;
      lrows_tmp.clear();
      for (int t = (int)out.size() - 1; t >= 0; --t) {
        int v = out[t];
        if (pinv[v] >= 0) continue;
        lrows_tmp.push_back(v);
      }
      for (size_t t = 0; t < lrows_tmp.size(); ++t) {
        int v = lrows_tmp[t];
        double a = sabs(x[v]);
        if (a > amax) { amax = a; piv = v; }
        if (v == diag_row[k]) dpos = v;
      }
      if (piv < 0) { ok = false; return false; }
      if (dpos >= 0 && sabs(x[dpos]) >= tol * amax) piv = dpos;
      T pv = x[piv];
      if (siszero(pv)) { ok = false; return false; }
      udiag[k] = pv;
      pinv[piv] = k;
      prow[k] = piv;
      x[piv] = szero<T>();
      for (int v : lrows_tmp) {
        if (v == piv) continue;
        li.push_back(v);
        lx.push_back(sdiv(x[v], pv));
        x[v] = szero<T>();
      }
      lp.push_back((int)li.size());
    }
    for (auto& r : li) r = pinv[r];  // L rows -> pivot steps
    ok = true;
    return true;
  }

  // Numeric refactor with the stored patterns and pivot order (no pivoting).
  bool refactor(const int* ap, const int* ai, const T* ax) {
    if (!ok) return false;
    // x is indexed by pivot step here
    for (int k = 0; k < n; ++k) {
      int col = q[k];
      for (int p = ap[col]; p < ap[col + 1]; ++p) x[pinv[ai[p]]] = ax[p];
      for (int p = up[k]; p < up[k + 1]; ++p) {
        int j = ui[p];
        T ujk = x[j];
        x[j] = szero<T>();
        ux[p] = ujk;
        for (int t = lp[j]; t < lp[j + 1]; ++t) x[li[t]] = ssub_mul(x[li[t]], lx[t], ujk);
      }
      T ukk = x[k];
      x[k] = szero<T>();
      if (siszero(ukk)) return false;
      udiag[k] = ukk;
      for (int t = lp[k]; t < lp[k + 1]; ++t) {
        lx[t] = sdiv(x[li[t]], ukk);
        x[li[t]] = szero<T>();
      }
    }
    return true;
  }

  // Solves A x = b; returns false when the result is not finite.
  bool solve(const T* b, T* out) {
    std::vector<T> y(n);
    for (int k = 0; k < n; ++k) y[k] = b[prow[k]];
    for (int k = 0; k < n; ++k) {
      T yk = y[k];
      for (int p = lp[k]; p < lp[k + 1]; ++p) y[li[p]] = ssub_mul(y[li[p]], lx[p], yk);
    }
    for (int k = n - 1; k >= 0; --k) {
      y[k] = sdiv(y[k], udiag[k]);
      T yk = y[k];
      for (int p = up[k]; p < up[k + 1]; ++p) y[ui[p]] = ssub_mul(y[ui[p]], ux[p], yk);
    }
    bool fin = true;
    for (int k = 0; k < n; ++k) {
      out[q[k]] = y[k];
      if (!sisfinite(y[k])) fin = false;
    }
    return fin;
  }
};

// Triplets (row, col, val) in emission order -> CSC with duplicates summed in emission order.
template <class T>
struct Csc {
  int n = 0;
  std::vector<int> ap, ai;
  std::vector<T> ax;
  std::vector<int> slot;  // triplet index -> CSC position
};
template <class T>
void build_csc(int n, const std::vector<int>& r, const std::vector<int>& c, Csc<T>& m) {
  size_t nt = r.size();
  std::vector<int> ord(nt);
  for (size_t i = 0; i < nt; ++i) ord[i] = (int)i;
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
    return c[a] != c[b] ? c[a] < c[b] : r[a] < r[b];
  });
  m.n = n;
  m.ap.assign(n + 1, 0);
  m.ai.clear();
  m.slot.assign(nt, -1);
  for (size_t t = 0; t < nt; ++t) {
    int e = ord[t];
    if (t > 0 && r[e] == r[ord[t - 1]] && c[e] == c[ord[t - 1]]) {
      m.slot[e] = (int)m.ai.size() - 1;
      continue;
    }
    m.ai.push_back(r[e]);
    m.ap[c[e] + 1]++;
    m.slot[e] = (int)m.ai.size() - 1;
  }
  for (int j = 0; j < n; ++j) m.ap[j + 1] += m.ap[j];
  m.ax.assign(m.ai.size(), szero<T>());
}
template <class T>
void pack_csc(Csc<T>& m, const T* vals, size_t nt) {
  std::fill(m.ax.begin(), m.ax.end(), szero<T>());
  for (size_t e = 0; e < nt; ++e) {
    T& d = m.ax[m.slot[e]];
    d = d + vals[e];
  }
}
inline cx operator+(cx a, cx b) { return cadd(a, b); }

// ---------------------------------------------------------------- network helpers
struct Net {
  int n;
  const int* rowptr;
  const int* col;
  const double* yre;
  const double* yim;
  const int* type;  // 0 slack, 1 PV, 2 PQ
  const double* zip = nullptr;  // per bus (P, Q) of constant-power, -current, -impedance ZIP terms, or null
  // optional elimination order of the Newton system's first factorization (unknowns in the
  // reference's numbering): column order colq[k] and preferred (diagonal) row rowp[k] of step k,
  // e.g. KLU's BTF + AMD orders; null = bus-major
  const int* colq = nullptr;
  const int* rowp = nullptr;
};

// effective_injection (src/network.rs:600-611) with the bus's ZIP terms summed per kind: the
// constant-power part, then S_p, S_i * vm and S_z * vm * vm added in that order.
void effective(const Net& net, int i, double vm, double& p, double& q) {
  if (!net.zip) return;
  const double* z = net.zip + 6 * (size_t)i;
  p = p + z[0];
  q = q + z[1];
  p = p + z[2] * vm;
  q = q + z[3] * vm;
  p = p + z[4] * vm * vm;
  q = q + z[5] * vm * vm;
}

void injections(const Net& net, const double* vm, const double* va, double* p, double* q) {
  std::vector<cx> v(net.n);
  for (int i = 0; i < net.n; ++i) v[i] = polar(vm[i], va[i]);
  for (int i = 0; i < net.n; ++i) {
    cx acc{0.0, 0.0};
    for (int t = net.rowptr[i]; t < net.rowptr[i + 1]; ++t)
      acc = cadd(acc, cmul({net.yre[t], net.yim[t]}, v[net.col[t]]));
    cx s = cmul(v[i], cconj(acc));
    p[i] = s.re;
    q[i] = s.im;
  }
}

struct Unknowns {
  std::vector<int> non_slack, pq, ns_pos, pq_pos;
  int n_ang = 0, n = 0;
};
Unknowns unknowns(const Net& net) {
  Unknowns u;
  u.ns_pos.assign(net.n, -1);
  u.pq_pos.assign(net.n, -1);
  for (int i = 0; i < net.n; ++i)
    if (net.type[i] != 0) { u.ns_pos[i] = (int)u.non_slack.size(); u.non_slack.push_back(i); }
  for (int i = 0; i < net.n; ++i)
    if (net.type[i] == 2) { u.pq_pos[i] = (int)u.pq.size(); u.pq.push_back(i); }
  u.n_ang = (int)u.non_slack.size();
  u.n = u.n_ang + (int)u.pq.size();
  return u;
}

// build_jacobian_triplets (src/solver.rs:818-888), same emission order and formulas.
void jacobian(const Net& net, const Unknowns& u, const double* vm, const double* va, const double* p,
              const double* q, std::vector<int>* rows, std::vector<int>* cols, std::vector<double>& vals) {
  vals.clear();
  if (rows) { rows->clear(); cols->clear(); }
  auto push = [&](int r, int c, double v) {
    if (rows) { rows->push_back(r); cols->push_back(c); }
    vals.push_back(v);
  };
  for (int ri = 0; ri < (int)u.non_slack.size(); ++ri) {
    int i = u.non_slack[ri];
    for (int t = net.rowptr[i]; t < net.rowptr[i + 1]; ++t) {
      int k = net.col[t];
      double g = net.yre[t], b = net.yim[t];
      if (k == i) {
        push(ri, ri, -q[i] - vm[i] * vm[i] * b);
        if (u.pq_pos[i] >= 0) push(ri, u.n_ang + u.pq_pos[i], p[i] / vm[i] + vm[i] * g);
        continue;
      }
      double ang = va[i] - va[k];
      if (u.ns_pos[k] >= 0) push(ri, u.ns_pos[k], vm[i] * vm[k] * (g * std::sin(ang) - b * std::cos(ang)));
      if (u.pq_pos[k] >= 0) push(ri, u.n_ang + u.pq_pos[k], vm[i] * (g * std::cos(ang) + b * std::sin(ang)));
    }
  }
  for (int ri = 0; ri < (int)u.pq.size(); ++ri) {
    int i = u.pq[ri];
    for (int t = net.rowptr[i]; t < net.rowptr[i + 1]; ++t) {
      int k = net.col[t];
      double g = net.yre[t], b = net.yim[t];
      if (k == i) {
        push(u.n_ang + ri, u.ns_pos[i], p[i] - vm[i] * vm[i] * g);
        push(u.n_ang + ri, u.n_ang + ri, q[i] / vm[i] - vm[i] * b);
        continue;
      }
      double ang = va[i] - va[k];
      if (u.ns_pos[k] >= 0)
        push(u.n_ang + ri, u.ns_pos[k], -vm[i] * vm[k] * (g * std::cos(ang) + b * std::sin(ang)));
      if (u.pq_pos[k] >= 0)
        push(u.n_ang + ri, u.n_ang + u.pq_pos[k], vm[i] * (g * std::sin(ang) - b * std::cos(ang)));
    }
  }
}

// Bus-major unknown order: the angle unknown of each non-slack bus followed by its magnitude
// unknown when the bus is PQ.  Used as the column order and the preferred (diagonal) rows.
std::vector<int> bus_major_order(const Net& net, const Unknowns& u) {
  std::vector<int> o;
  o.reserve(u.n);
  for (int i = 0; i < net.n; ++i) {
    if (u.ns_pos[i] >= 0) o.push_back(u.ns_pos[i]);
    if (u.pq_pos[i] >= 0) o.push_back(u.n_ang + u.pq_pos[i]);
  }
  return o;
}

// One Newton solve with a persistent factorization (PersistentSolver + newton_raphson_cached).
struct NrCache {
  bool has_pattern = false;
  Csc<double> csc;
  SparseLU<double> lu;
  bool has_lu = false;
};

enum { ST_CONVERGED = 0, ST_MAXITER = 1, ST_SINGULAR = 2 };

int nr_solve(const Net& net, const Unknowns& u, const double* p_spec, const double* q_spec, double* vm, double* va,
             double tol, int max_iter, double* history, int* iters, NrCache& cache) {
  std::vector<double> p(net.n), q(net.n), f(u.n), dx(u.n), vals;
  std::vector<int> rows, cols;
  int it = 0;
  for (; it < max_iter; ++it) {
    injections(net, vm, va, p.data(), q.data());
    int m = 0;
    for (int i : u.non_slack) {
      double pe = p_spec[i], qe = q_spec[i];
      effective(net, i, vm[i], pe, qe);
      f[m++] = pe - p[i];
    }
    for (int i : u.pq) {
      double pe = p_spec[i], qe = q_spec[i];
      effective(net, i, vm[i], pe, qe);
      f[m++] = qe - q[i];
    }
    double mx = 0.0;
    for (int t = 0; t < u.n; ++t) mx = std::fmax(mx, std::fabs(f[t]));  // NaN-ignoring, as f64::max
    if (history) history[it] = mx;
    if (mx < tol) { *iters = it + 1; return ST_CONVERGED; }
    jacobian(net, u, vm, va, p.data(), q.data(), cache.has_pattern ? nullptr : &rows,
             cache.has_pattern ? nullptr : &cols, vals);
    if (!cache.has_pattern) {
      build_csc(u.n, rows, cols, cache.csc);
      cache.has_pattern = true;
    }
    pack_csc(cache.csc, vals.data(), vals.size());
    if (!cache.has_lu) {
      // the first factorization fixes the pivot order; when it fails the cache stays empty and the
      // next solve factors its own first Jacobian (src/solver.rs:711-716)
      std::vector<int> order = net.colq ? std::vector<int>(net.colq, net.colq + u.n) : bus_major_order(net, u);
      std::vector<int> diag = net.rowp ? std::vector<int>(net.rowp, net.rowp + u.n) : order;
      if (cache.lu.factor(u.n, cache.csc.ap.data(), cache.csc.ai.data(), cache.csc.ax.data(), order, diag, 1e-3))
        cache.has_lu = true;
      if (!cache.has_lu) { *iters = it + 1; return ST_SINGULAR; }
    } else {
      if (!cache.lu.refactor(cache.csc.ap.data(), cache.csc.ai.data(), cache.csc.ax.data())) {
        *iters = it + 1;
        return ST_SINGULAR;
      }
    }
    if (!cache.lu.solve(f.data(), dx.data())) { *iters = it + 1; return ST_SINGULAR; }
    m = 0;
    for (int i : u.non_slack) va[i] += dx[m++];
    for (int i : u.pq) vm[i] += dx[m++];
  }
  *iters = max_iter;
  return ST_MAXITER;
}

// linear_initial_guess (src/network.rs:627-678).
int linear_guess(const Net& net, const double* p_spec, const double* q_spec, double* vm, double* va) {
  std::vector<int> unk, pos(net.n, -1);
  for (int i = 0; i < net.n; ++i)
    if (net.type[i] == 2) { pos[i] = (int)unk.size(); unk.push_back(i); }
  if (unk.empty()) return 0;
  int m = (int)unk.size();
  std::vector<int> tr, tc;
  std::vector<cx> tv;
  std::vector<cx> rhs(m, cx{0.0, 0.0});
  for (int r = 0; r < m; ++r) {
    int i = unk[r];
    double pe = p_spec[i], qe = q_spec[i];
    effective(net, i, vm[i], pe, qe);
    cx yl = cneg(cconj({pe, qe}));
    bool diag_seen = false;
    for (int t = net.rowptr[i]; t < net.rowptr[i + 1]; ++t) {
      int j = net.col[t];
      cx y{net.yre[t], net.yim[t]};
      if (j == i) { diag_seen = true; y = cadd(y, yl); }
      if (pos[j] >= 0) { tr.push_back(r); tc.push_back(pos[j]); tv.push_back(y); }
      else rhs[r] = csub(rhs[r], cmul(y, polar(vm[j], va[j])));
    }
    if (!diag_seen) { tr.push_back(r); tc.push_back(r); tv.push_back(yl); }
  }
  Csc<cx> a;
  build_csc(m, tr, tc, a);
  pack_csc(a, tv.data(), tv.size());
  std::vector<int> order(m);
  for (int k = 0; k < m; ++k) order[k] = k;
  SparseLU<cx> lu;
  std::vector<cx> sol(m);
  if (!lu.factor(m, a.ap.data(), a.ai.data(), a.ax.data(), order, order, 1e-3)) return 1;
  if (!lu.solve(rhs.data(), sol.data())) return 1;
  for (int r = 0; r < m; ++r) {
    vm[unk[r]] = cabs_(sol[r]);
    va[unk[r]] = std::atan2(sol[r].im, sol[r].re);
  }
  return 0;
}

}  // namespace

extern "C" {

// Builds the merged Y-bus in CSR form (rows sorted by column).  Transformer arrays carry complex
// values as interleaved (re, im) pairs.  Returns nnz, or -nnz_needed when cap is too small.
int orc_ybus(int n, int nl, const int* lf, const int* lt, const double* lr, const double* lx, const double* lb,
             const double* lg, int nt, const int* tf, const int* tt, const double* ys, const double* ysh,
             const double* tap, int cap, int* rowptr, int* col, double* yre, double* yim) {
  std::vector<std::pair<int, int>> key;
  std::vector<cx> val;
  auto add = [&](int i, int j, cx v) { key.push_back({i, j}); val.push_back(v); };
  for (int e = 0; e < nl; ++e) {
    if (lf[e] == lt[e]) { add(lf[e], lf[e], {lg[e], lb[e]}); continue; }
    cx yl = cdiv({1.0, 0.0}, {lr[e], lx[e]});
    cx ysh2{lg[e] / 2.0, lb[e] / 2.0};
    add(lf[e], lf[e], cadd(yl, ysh2));
    add(lt[e], lt[e], cadd(yl, ysh2));
    add(lf[e], lt[e], cneg(yl));
    add(lt[e], lf[e], cneg(yl));
  }
  for (int e = 0; e < nt; ++e) {
    cx s{ys[2 * e], ys[2 * e + 1]}, sh{ysh[2 * e], ysh[2 * e + 1]}, a{tap[2 * e], tap[2 * e + 1]};
    double k = cabs_(a);
    cx yd = cadd(s, cscale(sh, 0.5));
    add(tf[e], tf[e], cdiv(yd, {k * k, 0.0}));
    add(tf[e], tt[e], cdiv(cneg(s), cconj(a)));
    add(tt[e], tf[e], cdiv(cneg(s), a));
    add(tt[e], tt[e], yd);
  }
  std::map<std::pair<int, int>, cx> merged;
  for (size_t t = 0; t < key.size(); ++t) {
    auto it = merged.find(key[t]);
    if (it == merged.end()) merged[key[t]] = cadd({0.0, 0.0}, val[t]);
    else it->second = cadd(it->second, val[t]);
  }
  int nnz = (int)merged.size();
  if (nnz > cap) return -nnz;
  for (int i = 0; i <= n; ++i) rowptr[i] = 0;
  int t = 0;
  for (auto& kv : merged) {  // std::map order = (row, col) ascending
    rowptr[kv.first.first + 1]++;
    col[t] = kv.first.second;
    yre[t] = kv.second.re;
    yim[t] = kv.second.im;
    ++t;
  }
  for (int i = 0; i < n; ++i) rowptr[i + 1] += rowptr[i];
  return nnz;
}

void orc_power_injections(int n, const int* rowptr, const int* col, const double* yre, const double* yim,
                          const double* vm, const double* va, double* p, double* q) {
  Net net{n, rowptr, col, yre, yim, nullptr};
  injections(net, vm, va, p, q);
}

// Jacobian triplets in the reference's emission order.  Returns nnz or -nnz_needed.
int orc_jacobian(int n, const int* rowptr, const int* col, const double* yre, const double* yim, const int* type,
                 const double* vm, const double* va, const double* p, const double* q, int cap, int* rows,
                 int* cols, double* vals) {
  Net net{n, rowptr, col, yre, yim, type};
  Unknowns u = unknowns(net);
  std::vector<int> r, c;
  std::vector<double> v;
  jacobian(net, u, vm, va, p, q, &r, &c, v);
  int nnz = (int)v.size();
  if (nnz > cap) return -nnz;
  std::memcpy(rows, r.data(), sizeof(int) * nnz);
  std::memcpy(cols, c.data(), sizeof(int) * nnz);
  std::memcpy(vals, v.data(), sizeof(double) * nnz);
  return nnz;
}

int orc_linear_initial_guess(int n, const int* rowptr, const int* col, const double* yre, const double* yim,
                             const int* type, const double* p_spec, const double* q_spec, double* vm, double* va) {
  Net net{n, rowptr, col, yre, yim, type};
  return linear_guess(net, p_spec, q_spec, vm, va);
}

// orc_linear_initial_guess with ZIP terms (zip: 6 per bus as in Net, or null).
int orc_linear_initial_guess_zip(int n, const int* rowptr, const int* col, const double* yre, const double* yim,
                                 const int* type, const double* zip, const double* p_spec, const double* q_spec,
                                 double* vm, double* va) {
  Net net{n, rowptr, col, yre, yim, type, zip};
  return linear_guess(net, p_spec, q_spec, vm, va);
}

// Real sparse solve A x = b through the oracle LU (CSC input, bus-agnostic natural column order).
// Returns 0 on success, 1 when singular / non-finite.
int orc_sparse_solve(int n, const int* ap, const int* ai, const double* ax, const double* b, double* x) {
  SparseLU<double> lu;
  std::vector<int> order(n);
  for (int k = 0; k < n; ++k) order[k] = k;
  if (!lu.factor(n, ap, ai, ax, order, order, 1e-3)) return 1;
  return lu.solve(b, x) ? 0 : 1;
}

// Batch of independent NR solves on one topology (BatchSolver semantics with overrides applied
// by the caller).  Arrays are scenario-major [n_scen][n].  init: 0 = start from vm0/va0 as given,
// 1 = linear_initial_guess per scenario first.  Pivots are fixed by the first factorization each
// worker performs; with n_threads == 1 that is scenario 0's first Jacobian.
// Returns 0; writes status[s] (0 converged, 1 max iterations, 2 singular), iters[s] and
// history[s*max_iter + it].
int orc_batch_solve_zip(int n, const int* rowptr, const int* col, const double* yre, const double* yim,
                        const int* type, const double* zip, int n_scen, const double* p_spec, const double* q_spec,
                        const double* vm0, const double* va0, int init, double tol, int max_iter, int n_threads,
                        double* vm, double* va, int* status, int* iters, double* history);

int orc_batch_solve(int n, const int* rowptr, const int* col, const double* yre, const double* yim, const int* type,
                    int n_scen, const double* p_spec, const double* q_spec, const double* vm0, const double* va0,
                    int init, double tol, int max_iter, int n_threads, double* vm, double* va, int* status,
                    int* iters, double* history) {
  return orc_batch_solve_zip(n, rowptr, col, yre, yim, type, nullptr, n_scen, p_spec, q_spec, vm0, va0, init, tol,
                             max_iter, n_threads, vm, va, status, iters, history);
}

int orc_batch_solve_ordered(int n, const int* rowptr, const int* col, const double* yre, const double* yim,
                            const int* type, const double* zip, int n_scen, const double* p_spec,
                            const double* q_spec, const double* vm0, const double* va0, int init, double tol,
                            int max_iter, int n_threads, const int* colq, const int* rowp, double* vm, double* va,
                            int* status, int* iters, double* history);

// orc_batch_solve with ZIP terms (zip: 6 per bus as in Net, or null).
int orc_batch_solve_zip(int n, const int* rowptr, const int* col, const double* yre, const double* yim,
                        const int* type, const double* zip, int n_scen, const double* p_spec, const double* q_spec,
                        const double* vm0, const double* va0, int init, double tol, int max_iter, int n_threads,
                        double* vm, double* va, int* status, int* iters, double* history) {
  return orc_batch_solve_ordered(n, rowptr, col, yre, yim, type, zip, n_scen, p_spec, q_spec, vm0, va0, init, tol,
                                 max_iter, n_threads, nullptr, nullptr, vm, va, status, iters, history);
}

// orc_batch_solve_zip whose first factorizations take the elimination order colq / rowp (the
// unknowns' column order and preferred rows, as Net; null = bus-major).
int orc_batch_solve_ordered(int n, const int* rowptr, const int* col, const double* yre, const double* yim,
                            const int* type, const double* zip, int n_scen, const double* p_spec,
                            const double* q_spec, const double* vm0, const double* va0, int init, double tol,
                            int max_iter, int n_threads, const int* colq, const int* rowp, double* vm, double* va,
                            int* status, int* iters, double* history) {
  Net net{n, rowptr, col, yre, yim, type, zip, colq, rowp};
  Unknowns u = unknowns(net);
  std::atomic<int> next{0};
  auto worker = [&]() {
    NrCache cache;
    for (;;) {
      int s = next.fetch_add(1);
      if (s >= n_scen) break;
      double* v = vm + (size_t)s * n;
      double* a = va + (size_t)s * n;
      std::memcpy(v, vm0 + (size_t)s * n, sizeof(double) * n);
      std::memcpy(a, va0 + (size_t)s * n, sizeof(double) * n);
      const double* ps = p_spec + (size_t)s * n;
      const double* qs = q_spec + (size_t)s * n;
      if (init == 1) linear_guess(net, ps, qs, v, a);
      status[s] = nr_solve(net, u, ps, qs, v, a, tol, max_iter, history ? history + (size_t)s * max_iter : nullptr,
                           &iters[s], cache);
    }
  };
  if (n_threads <= 1) {
    worker();
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < n_threads; ++t) th.emplace_back(worker);
    for (auto& t : th) t.join();
  }
  return 0;
}

// Connected components of the Y-bus graph (src/network.rs:121-145): comp[i] = component id,
// components numbered by their lowest bus.  Returns the number of components.
int orc_components(int n, const int* rowptr, const int* col, int* comp) {
  for (int i = 0; i < n; ++i) comp[i] = -1;
  int nc = 0;
  std::vector<int> stack;
  for (int s = 0; s < n; ++s) {
    if (comp[s] >= 0) continue;
    comp[s] = nc;
    stack.assign(1, s);
    while (!stack.empty()) {
      int i = stack.back();
      stack.pop_back();
      for (int t = rowptr[i]; t < rowptr[i + 1]; ++t) {
        int j = col[t];
        if (j != i && comp[j] < 0) { comp[j] = nc; stack.push_back(j); }
      }
    }
    ++nc;
  }
  return nc;
}

}  // extern "C"
