// HIP kernels for the batched Newton-Raphson power flow on CDNA4 (gfx950).
//
// Data layout: scenarios are grouped in tiles of kTile (16); element e of scenario s lives at
//   base[(tile * count + e) * kTile + s % kTile],   tile = s / kTile,
// so the same element of the 16 scenarios of a tile is one contiguous 64 B (FP32) / 128 B (FP64)
// segment.  A 64-lane wavefront is split into kGroups = 4 lane groups of 16: lane (g, s) works on
// scenario s of the tile, and the 4 groups take 4 different elements (buses, L+U positions, rows
// of an L column) at once, so one vector access covers 4 x 16 contiguous values.
//
// Kernel  <- reference function it replaces
//   k_mismatch  power_injections + mismatch (src/network.rs:680-701, src/solver.rs:681-697)
//   k_decide    convergence test and history (src/solver.rs:697-703)
//   k_fill      JacobianPattern::fill_into (src/jacobian.rs:210-250), written straight into the
//               packed L+U storage at each entry's pivot position (pack_values_slice,
//               src/klu_native/mod.rs:123-125, plus the zero fill KLU's refactor scatters into)
//   k_lu        refactor_block_in_place (src/klu_native/kernel.rs:476-524) + solve
//               (src/klu_native/solve.rs:56-102): left-looking with the pivot order kept, column
//               k staged in LDS, forward substitution fused into the column loop
//   k_update    voltage update (src/solver.rs:722-732)
//   k_dv/k_resid  FP64 residual f - J*dx of the mixed-precision refinement sweep, evaluated from
//               the linearised injections dS = dV*conj(I) + V*conj(Y*dV) (the exact derivative
//               that the H/N/M/L formulas tabulate)
//
// Ownership rule inside k_lu (one wavefront per tile): every global L+U value is written and read
// by the lane group (position mod kGroups) of its scenario, and the right-hand side only by group
// 0, so no lane reads a global value that another lane of the same launch wrote; the one phase
// change that breaks this (back substitution reading U) is separated by a workgroup fence.
#include "gx_device.hpp"

#include <cmath>

namespace gx {
namespace {

constexpr int kGroups = kLanes / kTile;
// Mixed precision: stagnation below this mismatch (pu) is attributed to the FP32 factors.
constexpr double kStallFloor = 1e-4;
constexpr int kStride = kTile + 1;  // LDS row stride of a staged column (bank-conflict padding)

__device__ __forceinline__ size_t at(int tile, int count, int e) {
  return ((size_t)tile * (size_t)count + (size_t)e) * kTile;
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kLanes - 1); }
__device__ __forceinline__ int grp() { return lane_id() / kTile; }
__device__ __forceinline__ int sub() { return lane_id() % kTile; }

__device__ __forceinline__ bool tile_running(const int* status, int tile) {
  return __any(status[tile * kTile + sub()] == kRunning);
}

// true when v holds in every lane group of this lane's scenario
__device__ __forceinline__ bool all_groups(bool v) {
  int x = v ? 1 : 0;
  for (int d = kTile; d < kLanes; d <<= 1) x &= __shfl_xor(x, d);
  return x != 0;
}

// first index >= 0 of the arithmetic sequence base + i that lands on lane group g
__device__ __forceinline__ int first_of_group(int base, int g) {
  return (g - base % kGroups + kGroups) % kGroups;
}

// ---------------------------------------------------------------- layout conversion
// One block converts 64 elements of one tile: rows are read and written contiguously through LDS.
__global__ __launch_bounds__(256) void k_to_tiles(const double* __restrict__ src, double* __restrict__ dst,
                                                  int n_scen, int count) {
  __shared__ double buf[kTile][65];
  const int tile = blockIdx.y, e0 = blockIdx.x * 64;
  for (int idx = threadIdx.x; idx < kTile * 64; idx += blockDim.x) {
    const int r = idx / 64, c = idx % 64;
    int s = tile * kTile + r;
    s = s < n_scen ? s : n_scen - 1;  // padding scenarios repeat the last one
    const int e = e0 + c;
    buf[r][c] = e < count ? src[(size_t)s * count + e] : 0.0;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < kTile * 64; idx += blockDim.x) {
    const int c = idx / kTile, r = idx % kTile;
    const int e = e0 + c;
    if (e < count) dst[at(tile, count, e) + r] = buf[r][c];
  }
}

__global__ __launch_bounds__(256) void k_from_tiles(const double* __restrict__ src, double* __restrict__ dst,
                                                    int n_scen, int count) {
  __shared__ double buf[kTile][65];
  const int tile = blockIdx.y, e0 = blockIdx.x * 64;
  for (int idx = threadIdx.x; idx < kTile * 64; idx += blockDim.x) {
    const int c = idx / kTile, r = idx % kTile;
    const int e = e0 + c;
    buf[r][c] = e < count ? src[at(tile, count, e) + r] : 0.0;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < kTile * 64; idx += blockDim.x) {
    const int r = idx / 64, c = idx % 64;
    const int s = tile * kTile + r, e = e0 + c;
    if (s < n_scen && e < count) dst[(size_t)s * count + e] = buf[r][c];
  }
}

// Element-parallel kernels: grid (element blocks, tiles), 256 threads = 4 waves x 4 lane groups,
// so one block sweeps 16 elements (buses or L+U positions) at a time.
__device__ __forceinline__ int elem_first() {
  return (blockIdx.x * (blockDim.x / kLanes) + threadIdx.x / kLanes) * kGroups + grp();
}
__device__ __forceinline__ int elem_stride() { return gridDim.x * (blockDim.x / kLanes) * kGroups; }

__global__ __launch_bounds__(256) void k_init(DevTopo t, DevWork w) {
  const int tile = blockIdx.y, s = sub();
  if (blockIdx.x == 0 && threadIdx.x < kTile) {
    const int sc = tile * kTile + s;
    w.status[sc] = sc < w.n_scen ? kRunning : kPadLane;
    w.iters[sc] = 0;
    w.maxmis[sc] = 0ull;
  }
  if (blockIdx.x == 0 && tile == 0) {
    for (int it = threadIdx.x; it < w.max_iter; it += blockDim.x) w.running[it] = 0;
    if (threadIdx.x == 0 && w.fallback) *w.fallback = 0;
  }
  const size_t bb = at(tile, t.n_bus, 0) + s;
  for (int i = elem_first(); i < t.n_bus; i += elem_stride()) {
    const size_t o = bb + (size_t)i * kTile;
    if (t.bus_type[i] == kDeadBus) {  // member of an island without a Slack bus: V = 0, no injection
      w.vm[o] = 0.0;
      w.va[o] = 0.0;
      w.ps[o] = 0.0;
      w.qs[o] = 0.0;
    }
    const double m = w.vm[o], a = w.va[o];
    double sn, cs;
    sincos(a, &sn, &cs);
    w.vr[o] = m * cs;
    w.vi[o] = m * sn;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_mismatch(DevTopo t, DevWork w) {
  const int tile = blockIdx.y, s = sub();
  if (!tile_running(w.status, tile)) return;
  const int sc = tile * kTile + s;
  const size_t bb = at(tile, t.n_bus, 0) + s;
  const size_t ub = at(tile, t.n, 0) + s;
  T* __restrict__ y = reinterpret_cast<T*>(w.y);
  double mx = 0.0;
  for (int i = elem_first(); i < t.n_bus; i += elem_stride()) {
    double ir = 0.0, ii = 0.0;
    const int p2 = t.y_rowptr[i + 1];
    for (int p = t.y_rowptr[i]; p < p2; ++p) {
      const size_t oj = bb + (size_t)t.y_col[p] * kTile;
      const double g = t.y_g[p], b = t.y_b[p];
      const double vr = w.vr[oj], vi = w.vi[oj];
      ir += g * vr - b * vi;
      ii += g * vi + b * vr;
    }
    const size_t o = bb + (size_t)i * kTile;
    const double vr = w.vr[o], vi = w.vi[o];
    const double pcal = vr * ir + vi * ii;
    const double qcal = vi * ir - vr * ii;
    w.pc[o] = pcal;
    w.qc[o] = qcal;
    const int rp = t.rpos_p[i];
    if (rp >= 0) {
      const double fp = w.ps[o] - pcal;
      w.f[ub + (size_t)rp * kTile] = fp;
      y[ub + (size_t)rp * kTile] = (T)fp;
      mx = fmax(mx, fabs(fp));
    }
    const int rq = t.rpos_q[i];
    if (rq >= 0) {
      const double fq = w.qs[o] - qcal;
      w.f[ub + (size_t)rq * kTile] = fq;
      y[ub + (size_t)rq * kTile] = (T)fq;
      mx = fmax(mx, fabs(fq));
    }
  }
  // max over the lane groups of this scenario, then one atomic per scenario and wave
  for (int d = kTile; d < kLanes; d <<= 1) mx = fmax(mx, __shfl_xor(mx, d));
  if (grp() == 0 && w.status[sc] == kRunning && mx > 0.0)
    atomicMax(&w.maxmis[sc], (unsigned long long)__double_as_longlong(mx));
}

__global__ void k_decide(DevWork w, int iter) {
  const int sc = blockIdx.x * blockDim.x + threadIdx.x;
  if (sc >= w.n_tiles * kTile) return;
  if (w.status[sc] != kRunning) return;
  const double m = __longlong_as_double((long long)w.maxmis[sc]);
  w.maxmis[sc] = 0ull;
  w.history[(size_t)sc * w.max_iter + iter] = m;
  if (m < w.tol) {
    w.status[sc] = kConverged;
    w.iters[sc] = iter + 1;
  } else if (w.stall_check && iter >= 1 && w.history[(size_t)sc * w.max_iter + iter - 1] < kStallFloor &&
             !(m < w.history[(size_t)sc * w.max_iter + iter - 1])) {
    // mixed precision: the mismatch stopped falling although it was already small (the FP32
    // factors limit the step); leave the scenario to the FP64 re-solve
    w.status[sc] = kStalled;
    w.iters[sc] = iter + 1;
  } else {
    atomicAdd(&w.running[iter], 1);
  }
}

// FP64 re-solve, step 1: every scenario that did not converge gets its starting point back.
__global__ __launch_bounds__(256) void k_fallback_restore(DevTopo t, DevWork w) {
  const int tile = blockIdx.y, s = sub();
  const int sc = tile * kTile + s;
  const int st = w.status[sc];
  if (st != kStalled && st != kSingular) return;  // no barrier follows
  const size_t bb = at(tile, t.n_bus, 0) + s;
  for (int i = elem_first(); i < t.n_bus; i += elem_stride()) {
    const size_t o = bb + (size_t)i * kTile;
    const double m = w.vm0[o], a = w.va0[o];
    w.vm[o] = m;
    w.va[o] = a;
    double sn, cs;
    sincos(a, &sn, &cs);
    w.vr[o] = m * cs;
    w.vi[o] = m * sn;
  }
}

// FP64 re-solve, step 2 (after step 1 completed): those scenarios run again.
__global__ void k_fallback_reset(DevWork w) {
  const int sc = blockIdx.x * blockDim.x + threadIdx.x;
  if (sc == 0)
    for (int it = 0; it < w.max_iter; ++it) w.running[it] = 0;
  if (sc >= w.n_tiles * kTile) return;
  const int st = w.status[sc];
  if (st != kStalled && st != kSingular) return;
  w.status[sc] = kRunning;
  w.iters[sc] = 0;
  w.maxmis[sc] = 0ull;
  atomicAdd(w.fallback, 1);
}

// Jacobian values into the packed L+U storage, in a branch-free form of the eight formulas: with
// C = V_i V_k cos(t_i - t_k) and S = V_i V_k sin(t_i - t_k) (C = V_i^2, S = 0 on the diagonal),
//   H = G S - B C,  N = (G C + B S) / V_k,  M = -(G C + B S),  L = (G S - B C) / V_k,
// plus the injection term of the diagonal kinds: H -= Q_i, N += P_i / V_i, M += P_i,
// L += Q_i / V_i (src/jacobian.rs:220-243).
template <typename T>
__global__ __launch_bounds__(256) void k_fill(DevTopo t, DevWork w) {
  const int tile = blockIdx.y, s = sub();
  if (!tile_running(w.status, tile)) return;
  const size_t bb = at(tile, t.n_bus, 0) + s;
  T* __restrict__ lu = reinterpret_cast<T*>(w.lu) + at(tile, t.nnz_lu, 0) + s;
  for (int p = elem_first(); p < t.nnz_lu; p += elem_stride()) {
    const int kind = t.lu_kind[p];
    double v = 0.0;
    if (kind != kFill) {
      const int i = t.lu_bi[p], k = t.lu_bk[p];
      const double g = t.lu_g[p], b = t.lu_b[p];
      const size_t oi = bb + (size_t)i * kTile, ok = bb + (size_t)k * kTile;
      const double vri = w.vr[oi], vii = w.vi[oi], vrk = w.vr[ok], vik = w.vi[ok];
      const double cc = vri * vrk + vii * vik;
      const double ss = vii * vrk - vri * vik;
      const double hl = g * ss - b * cc;
      const double nm = g * cc + b * ss;
      const bool is_h = kind == kHii || kind == kHik;
      const bool is_n = kind == kNii || kind == kNik;
      const bool is_m = kind == kMii || kind == kMik;
      const bool diag = kind == kHii || kind == kNii || kind == kMii || kind == kLii;
      double base = is_h ? hl : (is_n ? nm : (is_m ? -nm : hl));
      const double vmk = w.vm[ok];
      if (is_n || !(is_h || is_m)) base = base / vmk;
      if (diag) {
        const double pc = w.pc[oi], qc = w.qc[oi];
        base += is_h ? -qc : (is_n ? pc / vmk : (is_m ? pc : qc / vmk));
      }
      v = base;
    }
    lu[(size_t)p * kTile] = (T)v;
  }
}

// Complex double for the linear initial guess, with num-complex 0.4 arithmetic (mul: (ac-bd,
// ad+bc); div: by the squared norm).
struct zd {
  double re, im;
};
__device__ __forceinline__ zd operator-(zd a, zd b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ zd operator*(zd a, zd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ zd operator/(zd a, zd b) {
  const double ns = b.re * b.re + b.im * b.im;
  return {(a.re * b.re + a.im * b.im) / ns, (a.im * b.re - a.re * b.im) / ns};
}
__device__ __forceinline__ bool is_zero(float v) { return v == 0.0f; }
__device__ __forceinline__ bool is_zero(double v) { return v == 0.0; }
__device__ __forceinline__ bool is_zero(zd v) { return v.re == 0.0 && v.im == 0.0; }
__device__ __forceinline__ bool finite_value(float v) { return isfinite(v); }
__device__ __forceinline__ bool finite_value(double v) { return isfinite(v); }
__device__ __forceinline__ bool finite_value(zd v) { return isfinite(v.re) && isfinite(v.im); }

// Refactor (optional) + forward/back substitution; one wavefront per tile of 16 scenarios.
// Column k is staged in LDS (x[o][s], o < len(k) <= cap), updated there by the 4 lane groups
// (distinct target rows within one update), divided by its pivot and written back.
template <typename T, bool kFactor>
__global__ __launch_bounds__(64) void k_lu(DevTopo t, DevWork w, int cap, int iter) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* x = reinterpret_cast<T*>(smem_raw);
  const int tile = blockIdx.x, g = grp(), s = sub();
  if (!tile_running(w.status, tile)) return;
  const int sc = tile * kTile + s;
  T* __restrict__ lu = reinterpret_cast<T*>(w.lu) + at(tile, t.nnz_lu, 0) + s;
  T* __restrict__ y = reinterpret_cast<T*>(w.y) + at(tile, t.n, 0) + s;
  bool ok = true;
  const int n = t.n;
  for (int k = 0; k < n; ++k) {
    const int c0 = t.cp[k], len = t.cp[k + 1] - c0, ul = t.ulen[k];
    if (kFactor) {
      if (len > cap) {  // the host sizes cap to the longest column; kept as a guard
        ok = false;
        continue;
      }
      const int o0 = first_of_group(c0, g);
      for (int o = o0; o < len; o += kGroups) x[o * kStride + s] = lu[(size_t)(c0 + o) * kTile];
      for (int u = 0; u < ul; ++u) {
        const T ujk = x[u * kStride + s];
        const int j = t.lu_row[c0 + u];
        const int l0 = t.cp[j] + t.ulen[j] + 1;
        const int ll = t.cp[j + 1] - l0;
        const uint16_t* __restrict__ off = t.upd_off + t.upd_ptr[c0 + u];
        int e = first_of_group(l0, g);
        // two rows per lane per step: both loads issued before the read-modify-writes
        for (; e + kGroups < ll; e += 2 * kGroups) {
          const T l1 = lu[(size_t)(l0 + e) * kTile];
          const T l2 = lu[(size_t)(l0 + e + kGroups) * kTile];
          const int t1 = off[e] * kStride + s, t2 = off[e + kGroups] * kStride + s;
          const T a1 = x[t1], a2 = x[t2];
          x[t1] = a1 - l1 * ujk;
          x[t2] = a2 - l2 * ujk;
        }
        if (e < ll) {
          const int t1 = off[e] * kStride + s;
          x[t1] = x[t1] - lu[(size_t)(l0 + e) * kTile] * ujk;
        }
      }
      const T ukk = x[ul * kStride + s];
      if (is_zero(ukk)) ok = false;
      for (int o = ul + 1 + first_of_group(ul + 1, g); o < len; o += kGroups) x[o * kStride + s] = x[o * kStride + s] / ukk;
      for (int o = o0; o < len; o += kGroups) lu[(size_t)(c0 + o) * kTile] = x[o * kStride + s];
      // forward substitution with column k of L (from LDS); the right-hand side belongs to group 0
      if (g == 0) {
        const T yk = y[(size_t)k * kTile];
        for (int o = ul + 1; o < len; ++o) {
          const size_t yi = (size_t)t.lu_row[c0 + o] * kTile;
          y[yi] = y[yi] - x[o * kStride + s] * yk;
        }
      }
    } else if (g == 0) {
      const T yk = y[(size_t)k * kTile];
      for (int o = ul + 1; o < len; ++o) {
        const size_t yi = (size_t)t.lu_row[c0 + o] * kTile;
        y[yi] = y[yi] - lu[(size_t)(c0 + o) * kTile] * yk;
      }
    }
  }
  if (kFactor) {
    // U values were written by every lane group; make them visible before group 0 reads them
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __syncthreads();
  }
  if (g == 0) {
    for (int k = n - 1; k >= 0; --k) {
      const int c0 = t.cp[k], ul = t.ulen[k];
      const T xk = y[(size_t)k * kTile] / lu[(size_t)(c0 + ul) * kTile];
      y[(size_t)k * kTile] = xk;
      if (!finite_value(xk)) ok = false;
      for (int u = 0; u < ul; ++u) {
        const size_t yj = (size_t)t.lu_row[c0 + u] * kTile;
        y[yj] = y[yj] - lu[(size_t)(c0 + u) * kTile] * xk;
      }
    }
  }
  // a zero pivot seen by any group of the scenario, or a non-finite solution, makes it Singular
  ok = all_groups(ok);
  if (g == 0 && !ok && w.status[sc] == kRunning) {
    w.status[sc] = kSingular;
    w.iters[sc] = iter + 1;
  }
}

// Mixed precision, refinement step 1: dx1 (FP64 copy of the FP32 solution) and
// dV_i = V_i * (d|V|_i / |V|_i + j dtheta_i) per bus.
template <typename T>
__global__ __launch_bounds__(256) void k_dv(DevTopo t, DevWork w, double* __restrict__ dvr,
                                            double* __restrict__ dvi) {
  const int tile = blockIdx.y, s = sub();
  if (!tile_running(w.status, tile)) return;
  const size_t bb = at(tile, t.n_bus, 0) + s;
  const size_t ub = at(tile, t.n, 0) + s;
  const T* __restrict__ y = reinterpret_cast<const T*>(w.y);
  for (int i = elem_first(); i < t.n_bus; i += elem_stride()) {
    const size_t o = bb + (size_t)i * kTile;
    double da = 0.0, dm = 0.0;
    const int xa = t.xpos_a[i], xm = t.xpos_m[i];
    if (xa >= 0) {
      da = (double)y[ub + (size_t)xa * kTile];
      w.dx[ub + (size_t)xa * kTile] = da;
    }
    if (xm >= 0) {
      dm = (double)y[ub + (size_t)xm * kTile];
      w.dx[ub + (size_t)xm * kTile] = dm;
    }
    const double a = dm / w.vm[o];
    const double vr = w.vr[o], vi = w.vi[o];
    dvr[o] = vr * a - vi * da;
    dvi[o] = vr * da + vi * a;
  }
}

// Mixed precision, refinement step 2: r = f - J*dx1 in FP64, stored (in LU precision) as the
// right-hand side of the correction solve.
template <typename T>
__global__ __launch_bounds__(256) void k_resid(DevTopo t, DevWork w, const double* __restrict__ dvr,
                                               const double* __restrict__ dvi) {
  const int tile = blockIdx.y, s = sub();
  if (!tile_running(w.status, tile)) return;
  const size_t bb = at(tile, t.n_bus, 0) + s;
  const size_t ub = at(tile, t.n, 0) + s;
  T* __restrict__ y = reinterpret_cast<T*>(w.y);
  for (int i = elem_first(); i < t.n_bus; i += elem_stride()) {
    const int rp = t.rpos_p[i], rq = t.rpos_q[i];
    if (rp < 0 && rq < 0) continue;
    double ar = 0.0, ai = 0.0;
    const int p2 = t.y_rowptr[i + 1];
    for (int p = t.y_rowptr[i]; p < p2; ++p) {
      const size_t oj = bb + (size_t)t.y_col[p] * kTile;
      const double g = t.y_g[p], b = t.y_b[p];
      const double dr = dvr[oj], di = dvi[oj];
      ar += g * dr - b * di;
      ai += g * di + b * dr;
    }
    const size_t o = bb + (size_t)i * kTile;
    const double vr = w.vr[o], vi = w.vi[o];
    const int xa = t.xpos_a[i], xm = t.xpos_m[i];
    const double bq = xa >= 0 ? w.dx[ub + (size_t)xa * kTile] : 0.0;
    const double aq = xm >= 0 ? w.dx[ub + (size_t)xm * kTile] / w.vm[o] : 0.0;
    const double pc = w.pc[o], qc = w.qc[o];
    const double dsr = (aq * pc - bq * qc) + (vr * ar + vi * ai);
    const double dsi = (aq * qc + bq * pc) + (vi * ar - vr * ai);
    if (rp >= 0) y[ub + (size_t)rp * kTile] = (T)(w.f[ub + (size_t)rp * kTile] - dsr);
    if (rq >= 0) y[ub + (size_t)rq * kTile] = (T)(w.f[ub + (size_t)rq * kTile] - dsi);
  }
}

template <typename T, bool kMixed>
__global__ __launch_bounds__(256) void k_update(DevTopo t, DevWork w) {
  const int tile = blockIdx.y, s = sub();
  if (!tile_running(w.status, tile)) return;
  const int sc = tile * kTile + s;
  if (w.status[sc] != kRunning) return;  // no barrier follows
  const size_t bb = at(tile, t.n_bus, 0) + s;
  const size_t ub = at(tile, t.n, 0) + s;
  const T* __restrict__ y = reinterpret_cast<const T*>(w.y);
  for (int i = elem_first(); i < t.n_bus; i += elem_stride()) {
    const int xa = t.xpos_a[i], xm = t.xpos_m[i];
    if (xa < 0 && xm < 0) continue;
    const size_t o = bb + (size_t)i * kTile;
    double a = w.va[o], m = w.vm[o];
    if (xa >= 0) {
      double d = (double)y[ub + (size_t)xa * kTile];
      if (kMixed) d += w.dx[ub + (size_t)xa * kTile];
      a += d;
    }
    if (xm >= 0) {
      double d = (double)y[ub + (size_t)xm * kTile];
      if (kMixed) d += w.dx[ub + (size_t)xm * kTile];
      m += d;
    }
    w.va[o] = a;
    w.vm[o] = m;
    double sn, cs;
    sincos(a, &sn, &cs);
    w.vr[o] = m * cs;
    w.vi[o] = m * sn;
  }
}

// Max |mismatch| of every island at the final voltages (resolve_pending, src/solver.rs:234-265):
// slack buses skipped, P mismatch for PV and PQ buses, Q mismatch for PQ buses.
__global__ __launch_bounds__(256) void k_island_max(DevTopo t, DevWork w, const int* __restrict__ comp, int n_comp,
                                                    unsigned long long* __restrict__ out) {
  const int tile = blockIdx.y, s = sub();
  const int sc = tile * kTile + s;
  const size_t bb = at(tile, t.n_bus, 0) + s;
  for (int i = elem_first(); i < t.n_bus; i += elem_stride()) {
    const int ty = t.bus_type[i];
    if (ty != 1 && ty != 2) continue;
    double ir = 0.0, ii = 0.0;
    const int p2 = t.y_rowptr[i + 1];
    for (int p = t.y_rowptr[i]; p < p2; ++p) {
      const size_t oj = bb + (size_t)t.y_col[p] * kTile;
      const double g = t.y_g[p], b = t.y_b[p];
      const double vr = w.vr[oj], vi = w.vi[oj];
      ir += g * vr - b * vi;
      ii += g * vi + b * vr;
    }
    const size_t o = bb + (size_t)i * kTile;
    const double vr = w.vr[o], vi = w.vi[o];
    double mx = fabs(w.ps[o] - (vr * ir + vi * ii));
    if (ty == 2) mx = fmax(mx, fabs(w.qs[o] - (vi * ir - vr * ii)));
    if (sc < w.n_scen && mx > 0.0)
      atomicMax(&out[(size_t)sc * n_comp + comp[i]], (unsigned long long)__double_as_longlong(mx));
  }
}

// ---------------------------------------------------------------- linear initial guess
// linear_initial_guess (src/network.rs:627-678): every PQ load becomes the constant admittance
// -conj(S) on its Y-bus diagonal and Y'_nn U_n = -Y'_ns U_s is solved for the PQ voltages; a
// singular or non-finite solve leaves the starting point unchanged.
__global__ __launch_bounds__(256) void k_lg_fill(DevTopo t, DevLg lg, DevWork w, zd* __restrict__ lu,
                                                 int* __restrict__ lg_status) {
  const int tile = blockIdx.y, s = sub();
  if (blockIdx.x == 0 && threadIdx.x < kTile) {
    const int sc = tile * kTile + s;
    lg_status[sc] = sc < w.n_scen ? kRunning : kPadLane;
  }
  const size_t bb = at(tile, t.n_bus, 0) + s;
  zd* out = lu + at(tile, lg.nnz_lu, 0) + s;
  for (int p = elem_first(); p < lg.nnz_lu; p += elem_stride()) {
    zd v{0.0, 0.0};
    const int e = lg.src[p];
    if (e >= 0) v = zd{t.y_g[e], t.y_b[e]};
    const int d = lg.dbus[p];
    if (d >= 0) {
      const size_t o = bb + (size_t)d * kTile;
      v.re += -w.ps[o];  // -conj(P + jQ) = -P + jQ
      v.im += w.qs[o];
    }
    out[(size_t)p * kTile] = v;
  }
}

__global__ __launch_bounds__(256) void k_lg_rhs(DevTopo t, DevLg lg, DevWork w, zd* __restrict__ y) {
  const int tile = blockIdx.y, s = sub();
  const size_t bb = at(tile, t.n_bus, 0) + s;
  zd* out = y + at(tile, lg.m, 0) + s;
  for (int r = elem_first(); r < lg.m; r += elem_stride()) {
    zd b{0.0, 0.0};
    for (int q = lg.rhs_ptr[r]; q < lg.rhs_ptr[r + 1]; ++q) {
      const int e = lg.rhs_e[q];
      const size_t oj = bb + (size_t)lg.rhs_j[q] * kTile;
      double sn, cs;
      sincos(w.va[oj], &sn, &cs);
      const zd u{w.vm[oj] * cs, w.vm[oj] * sn};
      b = b - zd{t.y_g[e], t.y_b[e]} * u;
    }
    out[(size_t)lg.row_piv[r] * kTile] = b;
  }
}

__global__ __launch_bounds__(256) void k_lg_apply(DevTopo t, DevLg lg, DevWork w, const zd* __restrict__ y,
                                                  const int* __restrict__ lg_status) {
  const int tile = blockIdx.y, s = sub();
  const int sc = tile * kTile + s;
  if (lg_status[sc] != kRunning) return;  // singular or padding: keep the starting point
  const size_t bb = at(tile, t.n_bus, 0) + s;
  const zd* in = y + at(tile, lg.m, 0) + s;
  for (int r = elem_first(); r < lg.m; r += elem_stride()) {
    const zd u = in[(size_t)lg.row_x[r] * kTile];
    const size_t o = bb + (size_t)lg.row_bus[r] * kTile;
    w.vm[o] = hypot(u.re, u.im);
    w.va[o] = atan2(u.im, u.re);
  }
}

__global__ void k_finish(DevWork w) {
  const int sc = blockIdx.x * blockDim.x + threadIdx.x;
  if (sc >= w.n_tiles * kTile) return;
  if (w.status[sc] == kRunning) {
    w.status[sc] = kMaxIter;
    w.iters[sc] = w.max_iter;
  }
}

// Copies a compact vector into / out of scenario 0 of a single-tile interleaved array.
__global__ void k_lane0(const double* __restrict__ src, double* __restrict__ dst, int count, int to_tile) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= count) return;
  if (to_tile) dst[(size_t)e * kTile] = src[e];
  else dst[e] = src[(size_t)e * kTile];
}

inline int elem_blocks(int count) {
  int b = (count + 16 * 4 - 1) / (16 * 4);  // about 4 elements per lane group
  return b < 1 ? 1 : (b > 1024 ? 1024 : b);
}

}  // namespace

void launch_to_tiles(const double* src, double* dst, int n_scen, int n_tiles, int count, hipStream_t stream) {
  if (count <= 0 || n_tiles <= 0) return;
  hipLaunchKernelGGL(k_to_tiles, dim3((count + 63) / 64, n_tiles), dim3(256), 0, stream, src, dst, n_scen, count);
}

void launch_from_tiles(const double* src, double* dst, int n_scen, int count, hipStream_t stream) {
  const int n_tiles = (n_scen + kTile - 1) / kTile;
  if (count <= 0 || n_tiles <= 0) return;
  hipLaunchKernelGGL(k_from_tiles, dim3((count + 63) / 64, n_tiles), dim3(256), 0, stream, src, dst, n_scen, count);
}

void launch_init(const DevTopo& t, const DevWork& w, hipStream_t stream) {
  hipLaunchKernelGGL(k_init, dim3(elem_blocks(t.n_bus), w.n_tiles), dim3(256), 0, stream, t, w);
}

void launch_mismatch(const DevTopo& t, const DevWork& w, bool fp32, hipStream_t stream) {
  dim3 grid(elem_blocks(t.n_bus), w.n_tiles);
  if (fp32) hipLaunchKernelGGL(k_mismatch<float>, grid, dim3(256), 0, stream, t, w);
  else hipLaunchKernelGGL(k_mismatch<double>, grid, dim3(256), 0, stream, t, w);
}

void launch_decide(const DevTopo& t, const DevWork& w, int iter, hipStream_t stream) {
  (void)t;
  const int threads = w.n_tiles * kTile;
  hipLaunchKernelGGL(k_decide, dim3((threads + 255) / 256), dim3(256), 0, stream, w, iter);
}

void launch_fill(const DevTopo& t, const DevWork& w, bool fp32, hipStream_t stream) {
  dim3 grid(elem_blocks(t.nnz_lu), w.n_tiles);
  if (fp32) hipLaunchKernelGGL(k_fill<float>, grid, dim3(256), 0, stream, t, w);
  else hipLaunchKernelGGL(k_fill<double>, grid, dim3(256), 0, stream, t, w);
}

size_t lu_shmem(const DevTopo& t, bool fp32) {
  const int cap = t.max_col < 1 ? 1 : t.max_col;
  return (size_t)cap * kStride * (fp32 ? sizeof(float) : sizeof(double));
}

void launch_lu(const DevTopo& t, const DevWork& w, bool fp32, bool factor, int iter, hipStream_t stream) {
  const int cap = t.max_col < 1 ? 1 : t.max_col;
  const size_t shmem = factor ? lu_shmem(t, fp32) : 0;
  dim3 grid(w.n_tiles);
  if (fp32) {
    if (factor) hipLaunchKernelGGL((k_lu<float, true>), grid, dim3(64), shmem, stream, t, w, cap, iter);
    else hipLaunchKernelGGL((k_lu<float, false>), grid, dim3(64), shmem, stream, t, w, cap, iter);
  } else {
    if (factor) hipLaunchKernelGGL((k_lu<double, true>), grid, dim3(64), shmem, stream, t, w, cap, iter);
    else hipLaunchKernelGGL((k_lu<double, false>), grid, dim3(64), shmem, stream, t, w, cap, iter);
  }
}

void launch_refine(const DevTopo& t, const DevWork& w, double* dvr, double* dvi, hipStream_t stream) {
  dim3 grid(elem_blocks(t.n_bus), w.n_tiles);
  hipLaunchKernelGGL(k_dv<float>, grid, dim3(256), 0, stream, t, w, dvr, dvi);
  hipLaunchKernelGGL(k_resid<float>, grid, dim3(256), 0, stream, t, w, dvr, dvi);
}

void launch_update(const DevTopo& t, const DevWork& w, bool fp32, bool mixed, hipStream_t stream) {
  dim3 grid(elem_blocks(t.n_bus), w.n_tiles);
  if (fp32) {
    if (mixed) hipLaunchKernelGGL((k_update<float, true>), grid, dim3(256), 0, stream, t, w);
    else hipLaunchKernelGGL((k_update<float, false>), grid, dim3(256), 0, stream, t, w);
  } else {
    hipLaunchKernelGGL((k_update<double, false>), grid, dim3(256), 0, stream, t, w);
  }
}

void launch_fallback(const DevTopo& t, const DevWork& w, hipStream_t stream) {
  hipLaunchKernelGGL(k_fallback_restore, dim3(elem_blocks(t.n_bus), w.n_tiles), dim3(256), 0, stream, t, w);
  const int threads = w.n_tiles * kTile;
  hipLaunchKernelGGL(k_fallback_reset, dim3((threads + 255) / 256), dim3(256), 0, stream, w);
}

void launch_finish(const DevWork& w, hipStream_t stream) {
  const int threads = w.n_tiles * kTile;
  hipLaunchKernelGGL(k_finish, dim3((threads + 255) / 256), dim3(256), 0, stream, w);
}

void launch_island_max(const DevTopo& t, const DevWork& w, const int* comp, int n_comp, unsigned long long* out,
                       hipStream_t stream) {
  hipLaunchKernelGGL(k_island_max, dim3(elem_blocks(t.n_bus), w.n_tiles), dim3(256), 0, stream, t, w, comp, n_comp,
                     out);
}

void launch_linear_guess(const DevTopo& t, const DevLg& lg, const DevTopo& lgt, const DevWork& w, void* lu,
                         void* y, int* lg_status, int* lg_iters, hipStream_t stream) {
  if (lg.m <= 0) return;
  hipLaunchKernelGGL(k_lg_fill, dim3(elem_blocks(lg.nnz_lu), w.n_tiles), dim3(256), 0, stream, t, lg, w,
                     reinterpret_cast<zd*>(lu), lg_status);
  hipLaunchKernelGGL(k_lg_rhs, dim3(elem_blocks(lg.m), w.n_tiles), dim3(256), 0, stream, t, lg, w,
                     reinterpret_cast<zd*>(y));
  DevWork lw = w;
  lw.lu = lu;
  lw.y = y;
  lw.status = lg_status;
  lw.iters = lg_iters;
  const int cap = lgt.max_col < 1 ? 1 : lgt.max_col;
  const size_t shmem = (size_t)cap * kStride * sizeof(zd);
  hipLaunchKernelGGL((k_lu<zd, true>), dim3(w.n_tiles), dim3(64), shmem, stream, lgt, lw, cap, 0);
  hipLaunchKernelGGL(k_lg_apply, dim3(elem_blocks(lg.m), w.n_tiles), dim3(256), 0, stream, t, lg, w,
                     reinterpret_cast<const zd*>(y), lg_status);
}

size_t lg_shmem(const DevTopo& lgt) {
  const int cap = lgt.max_col < 1 ? 1 : lgt.max_col;
  return (size_t)cap * kStride * 2 * sizeof(double);
}

void launch_lane0(const double* src, double* dst, int count, bool to_tile, hipStream_t stream) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_lane0, dim3((count + 255) / 256), dim3(256), 0, stream, src, dst, count, to_tile ? 1 : 0);
}

}  // namespace gx
