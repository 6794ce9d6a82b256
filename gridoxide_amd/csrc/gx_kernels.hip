// HIP kernels for the batched Newton-Raphson power flow on CDNA4 (gfx950).
//
// Data layout: scenarios are grouped in tiles of TW (4, 8 or 16; a template parameter of every
// kernel); element e of scenario s lives at
//   base[(tile * count + e) * TW + s % TW],   tile = s / TW,
// so the same element of a tile's scenarios is one contiguous segment.  A 64-lane wavefront is
// split into G = 64 / TW lane groups: lane (g, s) works on scenario s of the tile and the G groups
// take G different elements (buses, L+U positions, rows of an L column) at once, so one vector
// access covers G x TW contiguous values.
//
// Kernel  <- reference function it replaces
//   k_mismatch  power_injections + mismatch (src/network.rs:680-701, src/solver.rs:681-697)
//   k_decide    convergence test and history (src/solver.rs:697-703)
//   k_fill      JacobianPattern::fill_into (src/jacobian.rs:210-250), written straight into the
//               packed L+U storage at each entry's pivot position (pack_values_slice,
//               src/klu_native/mod.rs:123-125, plus the zero fill KLU's refactor scatters into)
//   k_lu        refactor_block_in_place (src/klu_native/kernel.rs:476-524) + solve
//               (src/klu_native/solve.rs:56-102): left-looking refactor with the pivot order kept
//               (column k staged in LDS, the L rows of the next update prefetched), forward
//               substitution fused into the column loop, then back substitution; the right-hand
//               side lives in an LDS ring window during both substitutions
//   k_update    voltage update (src/solver.rs:722-732)
//   k_dv/k_resid  FP64 residual f - J*dx of the mixed-precision refinement sweep, evaluated from
//               the linearised injections dS = dV*conj(I) + V*conj(Y*dV) (the exact derivative
//               that the H/N/M/L formulas tabulate)
//
// Visibility rule inside k_lu (one wavefront per tile): a global value is read by another lane of
// the same launch only after a workgroup fence separates the write from the read.  Every L+U
// value is therefore written and read by the lane group (position mod G) of its scenario, the
// right-hand side is written by group 0, and a fence every kRingGap columns (and between the two
// substitutions) covers the reads of older values by other lanes.
#include "gx_device.hpp"

#include <cmath>
#include <type_traits>

namespace gx {
namespace {

// Mixed precision: stagnation below this mismatch (pu) is attributed to the FP32 factors.
constexpr double kStallFloor = 1e-4;

template <int TW>
__device__ __forceinline__ size_t at(int tile, int count, int e) {
  return ((size_t)tile * (size_t)count + (size_t)e) * TW;
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kLanes - 1); }
template <int TW>
__device__ __forceinline__ int grp() { return lane_id() / TW; }
template <int TW>
__device__ __forceinline__ int sub() { return lane_id() % TW; }

template <int TW>
__device__ __forceinline__ bool tile_running(const int* status, int tile) {
  return __any(status[tile * TW + sub<TW>()] == kRunning);
}

// first index i >= 0 with (base + i) % G == g
template <int G>
__device__ __forceinline__ int first_of_group(int base, int g) {
  return (g - base % G + G) % G;
}

__device__ __forceinline__ void workgroup_fence() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __syncthreads();
}

// Complex double for the linear initial guess, with num-complex 0.4 arithmetic (mul: (ac-bd,
// ad+bc); div: by the squared norm).
struct zd {
  double re, im;
};
__device__ __forceinline__ zd operator-(zd a, zd b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ zd operator+(zd a, zd b) { return {a.re + b.re, a.im + b.im}; }
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
template <typename T>
__device__ __forceinline__ T zero_value() { return T(0); }
template <>
__device__ __forceinline__ zd zero_value<zd>() { return zd{0.0, 0.0}; }
__device__ __forceinline__ float shfl_xor_v(float v, int d) { return __shfl_xor(v, d); }
__device__ __forceinline__ double shfl_xor_v(double v, int d) { return __shfl_xor(v, d); }
__device__ __forceinline__ zd shfl_xor_v(zd v, int d) { return zd{__shfl_xor(v.re, d), __shfl_xor(v.im, d)}; }
__device__ __forceinline__ float shfl_v(float v, int l) { return __shfl(v, l); }
__device__ __forceinline__ double shfl_v(double v, int l) { return __shfl(v, l); }
__device__ __forceinline__ zd shfl_v(zd v, int l) { return zd{__shfl(v.re, l), __shfl(v.im, l)}; }

// sum over the lane groups of this lane's scenario
template <int TW, typename T>
__device__ __forceinline__ T group_sum(T v) {
  for (int d = TW; d < kLanes; d <<= 1) v = v + shfl_xor_v(v, d);
  return v;
}

template <int TW>
__device__ __forceinline__ bool all_groups(bool v) {
  int x = v ? 1 : 0;
  for (int d = TW; d < kLanes; d <<= 1) x &= __shfl_xor(x, d);
  return x != 0;
}

// ---------------------------------------------------------------- layout conversion
// One block converts 64 elements of one tile: rows are read and written contiguously through LDS.
template <int TW>
__global__ __launch_bounds__(256) void k_to_tiles(const double* __restrict__ src, double* __restrict__ dst,
                                                  int n_scen, int count) {
  __shared__ double buf[TW][65];
  const int tile = blockIdx.y, e0 = blockIdx.x * 64;
  for (int idx = threadIdx.x; idx < TW * 64; idx += blockDim.x) {
    const int r = idx / 64, c = idx % 64;
    int s = tile * TW + r;
    s = s < n_scen ? s : n_scen - 1;  // padding scenarios repeat the last one
    const int e = e0 + c;
    buf[r][c] = e < count ? src[(size_t)s * count + e] : 0.0;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < TW * 64; idx += blockDim.x) {
    const int c = idx / TW, r = idx % TW;
    const int e = e0 + c;
    if (e < count) dst[at<TW>(tile, count, e) + r] = buf[r][c];
  }
}

template <int TW>
__global__ __launch_bounds__(256) void k_from_tiles(const double* __restrict__ src, double* __restrict__ dst,
                                                    int n_scen, int count) {
  __shared__ double buf[TW][65];
  const int tile = blockIdx.y, e0 = blockIdx.x * 64;
  for (int idx = threadIdx.x; idx < TW * 64; idx += blockDim.x) {
    const int c = idx / TW, r = idx % TW;
    const int e = e0 + c;
    buf[r][c] = e < count ? src[at<TW>(tile, count, e) + r] : 0.0;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < TW * 64; idx += blockDim.x) {
    const int r = idx / 64, c = idx % 64;
    const int s = tile * TW + r, e = e0 + c;
    if (s < n_scen && e < count) dst[(size_t)s * count + e] = buf[r][c];
  }
}

// Element-parallel kernels: grid (element blocks, tiles), 256 threads = 4 waves x G lane groups.
template <int TW>
__device__ __forceinline__ int elem_first() {
  return (blockIdx.x * (blockDim.x / kLanes) + threadIdx.x / kLanes) * (kLanes / TW) + grp<TW>();
}
template <int TW>
__device__ __forceinline__ int elem_stride() {
  return gridDim.x * (blockDim.x / kLanes) * (kLanes / TW);
}

template <int TW>
__global__ __launch_bounds__(256) void k_init(DevTopo t, DevWork w) {
  const int tile = blockIdx.y, s = sub<TW>();
  if (blockIdx.x == 0 && threadIdx.x < TW) {
    const int sc = tile * TW + s;
    w.status[sc] = sc < w.n_scen ? kRunning : kPadLane;
    w.iters[sc] = 0;
    w.maxmis[sc] = 0ull;
  }
  if (blockIdx.x == 0 && tile == 0) {
    for (int it = threadIdx.x; it < w.max_iter; it += blockDim.x) w.running[it] = 0;
    if (threadIdx.x == 0 && w.fallback) *w.fallback = 0;
  }
  const size_t bb = at<TW>(tile, t.n_bus, 0) + s;
  for (int i = elem_first<TW>(); i < t.n_bus; i += elem_stride<TW>()) {
    const size_t o = bb + (size_t)i * TW;
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

template <typename T, int TW>
__global__ __launch_bounds__(256) void k_mismatch(DevTopo t, DevWork w) {
  const int tile = blockIdx.y, s = sub<TW>();
  if (!tile_running<TW>(w.status, tile)) return;
  const int sc = tile * TW + s;
  const size_t bb = at<TW>(tile, t.n_bus, 0) + s;
  const size_t ub = at<TW>(tile, t.n, 0) + s;
  T* __restrict__ y = reinterpret_cast<T*>(w.y);
  double mx = 0.0;
  for (int i = elem_first<TW>(); i < t.n_bus; i += elem_stride<TW>()) {
    double ir = 0.0, ii = 0.0;
    const int p2 = t.y_rowptr[i + 1];
    for (int p = t.y_rowptr[i]; p < p2; ++p) {
      const size_t oj = bb + (size_t)t.y_col[p] * TW;
      const double g = t.y_g[p], b = t.y_b[p];
      const double vr = w.vr[oj], vi = w.vi[oj];
      ir += g * vr - b * vi;
      ii += g * vi + b * vr;
    }
    const size_t o = bb + (size_t)i * TW;
    const double vr = w.vr[o], vi = w.vi[o];
    const double pcal = vr * ir + vi * ii;
    const double qcal = vi * ir - vr * ii;
    w.pc[o] = pcal;
    w.qc[o] = qcal;
    const int rp = t.rpos_p[i];
    if (rp >= 0) {
      const double fp = w.ps[o] - pcal;
      w.f[ub + (size_t)rp * TW] = fp;
      y[ub + (size_t)rp * TW] = (T)fp;
      mx = fmax(mx, fabs(fp));
    }
    const int rq = t.rpos_q[i];
    if (rq >= 0) {
      const double fq = w.qs[o] - qcal;
      w.f[ub + (size_t)rq * TW] = fq;
      y[ub + (size_t)rq * TW] = (T)fq;
      mx = fmax(mx, fabs(fq));
    }
  }
  // max over the lane groups of this scenario, then one atomic per scenario and wave
  for (int d = TW; d < kLanes; d <<= 1) mx = fmax(mx, __shfl_xor(mx, d));
  if (grp<TW>() == 0 && w.status[sc] == kRunning && mx > 0.0)
    atomicMax(&w.maxmis[sc], (unsigned long long)__double_as_longlong(mx));
}

__global__ void k_decide(DevWork w, int iter) {
  const int sc = blockIdx.x * blockDim.x + threadIdx.x;
  if (sc >= w.n_tiles * w.tile) return;
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

// FP64 re-solve, step 1: every stalled or FP32-singular scenario gets its starting point back.
template <int TW>
__global__ __launch_bounds__(256) void k_fallback_restore(DevTopo t, DevWork w) {
  const int tile = blockIdx.y, s = sub<TW>();
  const int sc = tile * TW + s;
  const int st = w.status[sc];
  if (st != kStalled && st != kSingular) return;  // no barrier follows
  const size_t bb = at<TW>(tile, t.n_bus, 0) + s;
  for (int i = elem_first<TW>(); i < t.n_bus; i += elem_stride<TW>()) {
    const size_t o = bb + (size_t)i * TW;
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
  if (sc >= w.n_tiles * w.tile) return;
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
template <typename T, int TW>
__global__ __launch_bounds__(256) void k_fill(DevTopo t, DevWork w) {
  const int tile = blockIdx.y, s = sub<TW>();
  if (!tile_running<TW>(w.status, tile)) return;
  const size_t bb = at<TW>(tile, t.n_bus, 0) + s;
  T* __restrict__ lu = reinterpret_cast<T*>(w.lu) + at<TW>(tile, t.nnz_lu, 0) + s;
  for (int p = elem_first<TW>(); p < t.nnz_lu; p += elem_stride<TW>()) {
    const int kind = t.lu_kind[p];
    double v = 0.0;
    if (kind != kFill) {
      const int i = t.lu_bi[p], k = t.lu_bk[p];
      const double g = t.lu_g[p], b = t.lu_b[p];
      const size_t oi = bb + (size_t)i * TW, ok = bb + (size_t)k * TW;
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
    lu[(size_t)p * TW] = (T)v;
  }
}

// Refactor (optional) + forward/back substitution; one wavefront per tile.
//   LDS: xc[cap][TW+1] (the column being factored) and ring[R][TW+1] (the right-hand side rows
//   inside the moving window; row r lives in slot r mod R, R = W + kRingGap + 1 a power of two).
// Forward: at column k the ring holds rows [k - B, k + W] (B = kRingGap); y_k = ring[k]; L(:,k)
// pushes into the ring rows i <= k + W; row k + W + 1 then enters the ring (taking the slot of
// row k - B) as f minus the contributions of its L entries farther than W, gathered at that
// moment.  Back substitution mirrors this with U: the ring holds rows [k - W, k + B].
template <typename T, int TW, bool kFactor>
__global__ __launch_bounds__(64) void k_lu(DevTopo t, DevWork w, int cap, int iter) {
  constexpr int G = kLanes / TW;
  constexpr int S = TW + 1;
  constexpr int B = kRingGap;
  const int W = t.window;
  const int R = W + B + 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* xc = reinterpret_cast<T*>(smem_raw);
  T* ring = xc + (kFactor ? (size_t)cap * S : 0);
  const int tile = blockIdx.x, g = grp<TW>(), s = sub<TW>();
  if (!tile_running<TW>(w.status, tile)) return;
  const int sc = tile * TW + s;
  T* __restrict__ lu = reinterpret_cast<T*>(w.lu) + at<TW>(tile, t.nnz_lu, 0) + s;
  T* __restrict__ y = reinterpret_cast<T*>(w.y) + at<TW>(tile, t.n, 0) + s;
  const int n = t.n;
  bool ok = true;        // per lane: no zero pivot, finite solution
  bool overflow = false; // a column longer than the staging area (the same for every lane)
  auto slot = [&](int r) { return (r & (R - 1)) * S + s; };

  // ------------------------------------------------ forward pass (+ refactor)
  for (int r = g; r <= W && r < n; r += G) ring[slot(r)] = y[(size_t)r * TW];
  // f of the rows entering the window: within a batch of G columns, group g holds the row that
  // enters after column (batch start + g); the next batch is loaded one batch ahead
  auto f_of = [&](int kk) -> T {
    const int r = kk + W + 1;
    return r < n ? y[(size_t)r * TW] : zero_value<T>();
  };
  T fcur = f_of(g), fnext = f_of(G + g);
  for (int k = 0; k < n; ++k) {
    if (k > 0 && k % G == 0) {
      fcur = fnext;
      fnext = f_of(k + G + g);
    }
    const int c0 = t.cp[k], len = t.cp[k + 1] - c0, ul = t.ulen[k];
    if (kFactor) {
      if (len > cap) {  // the host sizes cap to the longest column; kept as a guard
        overflow = true;
        break;
      }
      for (int o = first_of_group<G>(c0, g); o < len; o += G) xc[o * S + s] = lu[(size_t)(c0 + o) * TW];
      if (ul > 0) {
        // updates by the columns j of U(:,k), in ascending order; the L rows of the next update
        // are loaded while the current one runs
        const int* __restrict__ rp = t.rec + 4 * (size_t)c0;
        int4 rc = *reinterpret_cast<const int4*>(rp);
        int e = first_of_group<G>(rc.x, g);
        bool pv = e < rc.y;
        T pl = pv ? lu[(size_t)(rc.x + e) * TW] : zero_value<T>();
        int po = pv ? t.upd_off[rc.z + e] : 0;
        for (int u = 0; u < ul; ++u) {
          const int4 cur = rc;
          const T cl = pl;
          const int co = po, ce = e;
          const bool cv = pv;
          if (u + 1 < ul) {
            rc = *reinterpret_cast<const int4*>(rp + 4 * (u + 1));
            e = first_of_group<G>(rc.x, g);
            pv = e < rc.y;
            pl = pv ? lu[(size_t)(rc.x + e) * TW] : zero_value<T>();
            po = pv ? t.upd_off[rc.z + e] : 0;
          }
          const T ujk = xc[u * S + s];
          if (cv) xc[co * S + s] = xc[co * S + s] - cl * ujk;
          // longer L columns: further rows, four per lane at a time (distinct target rows)
          int ee = ce + G;
          for (; ee + 3 * G < cur.y; ee += 4 * G) {
            T lv[4];
            int oo[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              lv[q] = lu[(size_t)(cur.x + ee + q * G) * TW];
              oo[q] = t.upd_off[cur.z + ee + q * G];
            }
            T xv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) xv[q] = xc[oo[q] * S + s];
#pragma unroll
            for (int q = 0; q < 4; ++q) xc[oo[q] * S + s] = xv[q] - lv[q] * ujk;
          }
          for (; ee < cur.y; ee += G) {
            const int oo = t.upd_off[cur.z + ee];
            xc[oo * S + s] = xc[oo * S + s] - lu[(size_t)(cur.x + ee) * TW] * ujk;
          }
        }
      }
      const T ukk = xc[ul * S + s];
      if (is_zero(ukk)) ok = false;
      for (int o = ul + 1 + first_of_group<G>(ul + 1, g); o < len; o += G) xc[o * S + s] = xc[o * S + s] / ukk;
      for (int o = first_of_group<G>(c0, g); o < len; o += G) lu[(size_t)(c0 + o) * TW] = xc[o * S + s];
    }
    // forward substitution: y_k is final; push L(:,k) into the window rows
    const T yk = ring[slot(k)];
    for (int o = ul + 1 + g; o < len; o += G) {
      const int i = t.lu_row[c0 + o];
      if (i - k > W) continue;  // gathered when row i enters the window
      const T lv = kFactor ? xc[o * S + s] : lu[(size_t)(c0 + o) * TW];
      ring[slot(i)] = ring[slot(i)] - lv * yk;
    }
    if (g == 0) y[(size_t)k * TW] = yk;
    const int r = k + W + 1;
    if (r < n) {
      T acc = zero_value<T>();
      for (int q = t.fwd_ptr[r]; q < t.fwd_ptr[r + 1]; ++q) {
        const int p = t.fwd_pos[q];
        if (p % G != g) continue;  // the lane group that wrote L at p reads it
        const int j = t.fwd_col[q];
        const T yj = j > k - B ? ring[slot(j)] : y[(size_t)j * TW];
        acc = acc + lu[(size_t)p * TW] * yj;
      }
      acc = group_sum<TW>(acc);
      if (g == k % G) ring[slot(r)] = fcur - acc;
    }
    if ((k + 1) % B == 0) workgroup_fence();
  }
  workgroup_fence();

  // ------------------------------------------------ back substitution
  if (!overflow) {
    // ring rows [max(0, n-1-W), n-1] <- y
    for (int r = n - 1 - g; r >= 0 && r >= n - 1 - W; r -= G) ring[slot(r)] = y[(size_t)r * TW];
    // within a batch of G columns (counted from the end), group g holds y of the row entering
    // after column n-1-(batch start + g) and that column's pivot
    auto y_of = [&](int m) -> T {
      const int r = n - 1 - m - W - 1;
      return r >= 0 ? y[(size_t)r * TW] : zero_value<T>();
    };
    auto piv_of = [&](int m) -> T {
      const int k = n - 1 - m;
      return k >= 0 ? lu[(size_t)(t.cp[k] + t.ulen[k]) * TW] : zero_value<T>();
    };
    T ycur = y_of(g), ynext = y_of(G + g);
    T pcur = piv_of(g), pnext = piv_of(G + g);
    for (int k = n - 1; k >= 0; --k) {
      const int m = n - 1 - k;
      if (m > 0 && m % G == 0) {
        ycur = ynext;
        ynext = y_of(m + G + g);
        pcur = pnext;
        pnext = piv_of(m + G + g);
      }
      const T ukk = shfl_v(pcur, (m % G) * TW + s);
      const T xk = ring[slot(k)] / ukk;
      if (!finite_value(xk)) ok = false;
      if (g == 0) {
        ring[slot(k)] = xk;
        y[(size_t)k * TW] = xk;
      }
      const int c0 = t.cp[k], ul = t.ulen[k];
      for (int u = g; u < ul; u += G) {
        const int j = t.lu_row[c0 + u];
        if (k - j > W) continue;  // gathered when row j enters the window
        ring[slot(j)] = ring[slot(j)] - lu[(size_t)(c0 + u) * TW] * xk;
      }
      const int r = k - W - 1;
      if (r >= 0) {
        T acc = zero_value<T>();
        for (int q = t.bwd_ptr[r] + g; q < t.bwd_ptr[r + 1]; q += G) {
          const int j = t.bwd_col[q];
          const T xj = j < k + B ? ring[slot(j)] : y[(size_t)j * TW];
          acc = acc + lu[(size_t)t.bwd_pos[q] * TW] * xj;
        }
        acc = group_sum<TW>(acc);
        if (g == m % G) ring[slot(r)] = ycur - acc;
      }
      if ((m + 1) % B == 0) workgroup_fence();
    }
  }
  // a zero pivot seen by any group of the scenario, or a non-finite solution, makes it Singular
  ok = all_groups<TW>(ok && !overflow);
  if (g == 0 && !ok && w.status[sc] == kRunning) {
    w.status[sc] = kSingular;
    w.iters[sc] = iter + 1;
  }
}

// Mixed precision, refinement step 1: dx1 (FP64 copy of the FP32 solution) and
// dV_i = V_i * (d|V|_i / |V|_i + j dtheta_i) per bus.
template <typename T, int TW>
__global__ __launch_bounds__(256) void k_dv(DevTopo t, DevWork w, double* __restrict__ dvr,
                                            double* __restrict__ dvi) {
  const int tile = blockIdx.y, s = sub<TW>();
  if (!tile_running<TW>(w.status, tile)) return;
  const size_t bb = at<TW>(tile, t.n_bus, 0) + s;
  const size_t ub = at<TW>(tile, t.n, 0) + s;
  const T* __restrict__ y = reinterpret_cast<const T*>(w.y);
  for (int i = elem_first<TW>(); i < t.n_bus; i += elem_stride<TW>()) {
    const size_t o = bb + (size_t)i * TW;
    double da = 0.0, dm = 0.0;
    const int xa = t.xpos_a[i], xm = t.xpos_m[i];
    if (xa >= 0) {
      da = (double)y[ub + (size_t)xa * TW];
      w.dx[ub + (size_t)xa * TW] = da;
    }
    if (xm >= 0) {
      dm = (double)y[ub + (size_t)xm * TW];
      w.dx[ub + (size_t)xm * TW] = dm;
    }
    const double a = dm / w.vm[o];
    const double vr = w.vr[o], vi = w.vi[o];
    dvr[o] = vr * a - vi * da;
    dvi[o] = vr * da + vi * a;
  }
}

// Mixed precision, refinement step 2: r = f - J*dx1 in FP64, stored (in LU precision) as the
// right-hand side of the correction solve.
template <typename T, int TW>
__global__ __launch_bounds__(256) void k_resid(DevTopo t, DevWork w, const double* __restrict__ dvr,
                                               const double* __restrict__ dvi) {
  const int tile = blockIdx.y, s = sub<TW>();
  if (!tile_running<TW>(w.status, tile)) return;
  const size_t bb = at<TW>(tile, t.n_bus, 0) + s;
  const size_t ub = at<TW>(tile, t.n, 0) + s;
  T* __restrict__ y = reinterpret_cast<T*>(w.y);
  for (int i = elem_first<TW>(); i < t.n_bus; i += elem_stride<TW>()) {
    const int rp = t.rpos_p[i], rq = t.rpos_q[i];
    if (rp < 0 && rq < 0) continue;
    double ar = 0.0, ai = 0.0;
    const int p2 = t.y_rowptr[i + 1];
    for (int p = t.y_rowptr[i]; p < p2; ++p) {
      const size_t oj = bb + (size_t)t.y_col[p] * TW;
      const double g = t.y_g[p], b = t.y_b[p];
      const double dr = dvr[oj], di = dvi[oj];
      ar += g * dr - b * di;
      ai += g * di + b * dr;
    }
    const size_t o = bb + (size_t)i * TW;
    const double vr = w.vr[o], vi = w.vi[o];
    const int xa = t.xpos_a[i], xm = t.xpos_m[i];
    const double bq = xa >= 0 ? w.dx[ub + (size_t)xa * TW] : 0.0;
    const double aq = xm >= 0 ? w.dx[ub + (size_t)xm * TW] / w.vm[o] : 0.0;
    const double pc = w.pc[o], qc = w.qc[o];
    const double dsr = (aq * pc - bq * qc) + (vr * ar + vi * ai);
    const double dsi = (aq * qc + bq * pc) + (vi * ar - vr * ai);
    if (rp >= 0) y[ub + (size_t)rp * TW] = (T)(w.f[ub + (size_t)rp * TW] - dsr);
    if (rq >= 0) y[ub + (size_t)rq * TW] = (T)(w.f[ub + (size_t)rq * TW] - dsi);
  }
}

template <typename T, int TW, bool kMixed>
__global__ __launch_bounds__(256) void k_update(DevTopo t, DevWork w) {
  const int tile = blockIdx.y, s = sub<TW>();
  if (!tile_running<TW>(w.status, tile)) return;
  const int sc = tile * TW + s;
  if (w.status[sc] != kRunning) return;  // no barrier follows
  const size_t bb = at<TW>(tile, t.n_bus, 0) + s;
  const size_t ub = at<TW>(tile, t.n, 0) + s;
  const T* __restrict__ y = reinterpret_cast<const T*>(w.y);
  for (int i = elem_first<TW>(); i < t.n_bus; i += elem_stride<TW>()) {
    const int xa = t.xpos_a[i], xm = t.xpos_m[i];
    if (xa < 0 && xm < 0) continue;
    const size_t o = bb + (size_t)i * TW;
    double a = w.va[o], m = w.vm[o];
    if (xa >= 0) {
      double d = (double)y[ub + (size_t)xa * TW];
      if (kMixed) d += w.dx[ub + (size_t)xa * TW];
      a += d;
    }
    if (xm >= 0) {
      double d = (double)y[ub + (size_t)xm * TW];
      if (kMixed) d += w.dx[ub + (size_t)xm * TW];
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
template <int TW>
__global__ __launch_bounds__(256) void k_island_max(DevTopo t, DevWork w, const int* __restrict__ comp, int n_comp,
                                                    unsigned long long* __restrict__ out) {
  const int tile = blockIdx.y, s = sub<TW>();
  const int sc = tile * TW + s;
  const size_t bb = at<TW>(tile, t.n_bus, 0) + s;
  for (int i = elem_first<TW>(); i < t.n_bus; i += elem_stride<TW>()) {
    const int ty = t.bus_type[i];
    if (ty != 1 && ty != 2) continue;
    double ir = 0.0, ii = 0.0;
    const int p2 = t.y_rowptr[i + 1];
    for (int p = t.y_rowptr[i]; p < p2; ++p) {
      const size_t oj = bb + (size_t)t.y_col[p] * TW;
      const double g = t.y_g[p], b = t.y_b[p];
      const double vr = w.vr[oj], vi = w.vi[oj];
      ir += g * vr - b * vi;
      ii += g * vi + b * vr;
    }
    const size_t o = bb + (size_t)i * TW;
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
template <int TW>
__global__ __launch_bounds__(256) void k_lg_fill(DevTopo t, DevLg lg, DevWork w, zd* __restrict__ lu,
                                                 int* __restrict__ lg_status) {
  const int tile = blockIdx.y, s = sub<TW>();
  if (blockIdx.x == 0 && threadIdx.x < TW) {
    const int sc = tile * TW + s;
    lg_status[sc] = sc < w.n_scen ? kRunning : kPadLane;
  }
  const size_t bb = at<TW>(tile, t.n_bus, 0) + s;
  zd* out = lu + at<TW>(tile, lg.nnz_lu, 0) + s;
  for (int p = elem_first<TW>(); p < lg.nnz_lu; p += elem_stride<TW>()) {
    zd v{0.0, 0.0};
    const int e = lg.src[p];
    if (e >= 0) v = zd{t.y_g[e], t.y_b[e]};
    const int d = lg.dbus[p];
    if (d >= 0) {
      const size_t o = bb + (size_t)d * TW;
      v.re += -w.ps[o];  // -conj(P + jQ) = -P + jQ
      v.im += w.qs[o];
    }
    out[(size_t)p * TW] = v;
  }
}

template <int TW>
__global__ __launch_bounds__(256) void k_lg_rhs(DevTopo t, DevLg lg, DevWork w, zd* __restrict__ y) {
  const int tile = blockIdx.y, s = sub<TW>();
  const size_t bb = at<TW>(tile, t.n_bus, 0) + s;
  zd* out = y + at<TW>(tile, lg.m, 0) + s;
  for (int r = elem_first<TW>(); r < lg.m; r += elem_stride<TW>()) {
    zd b{0.0, 0.0};
    for (int q = lg.rhs_ptr[r]; q < lg.rhs_ptr[r + 1]; ++q) {
      const int e = lg.rhs_e[q];
      const size_t oj = bb + (size_t)lg.rhs_j[q] * TW;
      double sn, cs;
      sincos(w.va[oj], &sn, &cs);
      const zd u{w.vm[oj] * cs, w.vm[oj] * sn};
      b = b - zd{t.y_g[e], t.y_b[e]} * u;
    }
    out[(size_t)lg.row_piv[r] * TW] = b;
  }
}

template <int TW>
__global__ __launch_bounds__(256) void k_lg_apply(DevTopo t, DevLg lg, DevWork w, const zd* __restrict__ y,
                                                  const int* __restrict__ lg_status) {
  const int tile = blockIdx.y, s = sub<TW>();
  const int sc = tile * TW + s;
  if (lg_status[sc] != kRunning) return;  // singular or padding: keep the starting point
  const size_t bb = at<TW>(tile, t.n_bus, 0) + s;
  const zd* in = y + at<TW>(tile, lg.m, 0) + s;
  for (int r = elem_first<TW>(); r < lg.m; r += elem_stride<TW>()) {
    const zd u = in[(size_t)lg.row_x[r] * TW];
    const size_t o = bb + (size_t)lg.row_bus[r] * TW;
    w.vm[o] = hypot(u.re, u.im);
    w.va[o] = atan2(u.im, u.re);
  }
}

__global__ void k_finish(DevWork w) {
  const int sc = blockIdx.x * blockDim.x + threadIdx.x;
  if (sc >= w.n_tiles * w.tile) return;
  if (w.status[sc] == kRunning) {
    w.status[sc] = kMaxIter;
    w.iters[sc] = w.max_iter;
  }
}

// Copies a compact vector into / out of scenario 0 of a single-tile interleaved array.
__global__ void k_lane0(const double* __restrict__ src, double* __restrict__ dst, int count, int to_tile, int tw) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= count) return;
  if (to_tile) dst[(size_t)e * tw] = src[e];
  else dst[e] = src[(size_t)e * tw];
}

inline int elem_blocks(int count, int tile) {
  const int per_block = 4 * (kLanes / tile) * 4;  // about 4 elements per lane group
  int b = (count + per_block - 1) / per_block;
  return b < 1 ? 1 : (b > 1024 ? 1024 : b);
}

// Tile-width dispatch: f is a generic lambda taking std::integral_constant<int, TW>.
template <class F>
void with_tile(int tile, F&& f) {
  switch (tile) {
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 8: f(std::integral_constant<int, 8>{}); break;
    default: f(std::integral_constant<int, 16>{}); break;
  }
}

template <typename T>
size_t lu_bytes(const DevTopo& t, int tile, bool factor) {
  const size_t cap = t.max_col < 1 ? 1 : t.max_col;
  const size_t rows = (size_t)t.window + kRingGap + 1;
  return ((factor ? cap : 0) + rows) * (size_t)(tile + 1) * sizeof(T);
}

// Raises a kernel's dynamic-LDS limit to what this launch needs (once per larger size).
template <class K>
void allow_lds(K kernel, size_t bytes, size_t& granted) {
  if (bytes <= granted) return;
  if (hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) == hipSuccess)
    granted = bytes;
}

template <typename T, int TW>
void run_lu(const DevTopo& t, const DevWork& w, bool factor, int iter, hipStream_t stream) {
  static size_t granted_f = 65536, granted_s = 65536;
  const int cap = t.max_col < 1 ? 1 : t.max_col;
  const size_t shmem = lu_bytes<T>(t, TW, factor);
  if (factor) {
    allow_lds(k_lu<T, TW, true>, shmem, granted_f);
    hipLaunchKernelGGL((k_lu<T, TW, true>), dim3(w.n_tiles), dim3(64), shmem, stream, t, w, cap, iter);
  } else {
    allow_lds(k_lu<T, TW, false>, shmem, granted_s);
    hipLaunchKernelGGL((k_lu<T, TW, false>), dim3(w.n_tiles), dim3(64), shmem, stream, t, w, cap, iter);
  }
}

}  // namespace

void launch_to_tiles(const double* src, double* dst, int n_scen, int n_tiles, int count, int tile,
                     hipStream_t stream) {
  if (count <= 0 || n_tiles <= 0) return;
  with_tile(tile, [&](auto tw) {
    hipLaunchKernelGGL(k_to_tiles<decltype(tw)::value>, dim3((count + 63) / 64, n_tiles), dim3(256), 0, stream, src,
                       dst, n_scen, count);
  });
}

void launch_from_tiles(const double* src, double* dst, int n_scen, int count, int tile, hipStream_t stream) {
  const int n_tiles = (n_scen + tile - 1) / tile;
  if (count <= 0 || n_tiles <= 0) return;
  with_tile(tile, [&](auto tw) {
    hipLaunchKernelGGL(k_from_tiles<decltype(tw)::value>, dim3((count + 63) / 64, n_tiles), dim3(256), 0, stream,
                       src, dst, n_scen, count);
  });
}

void launch_init(const DevTopo& t, const DevWork& w, hipStream_t stream) {
  with_tile(w.tile, [&](auto tw) {
    constexpr int TW = decltype(tw)::value;
    hipLaunchKernelGGL(k_init<TW>, dim3(elem_blocks(t.n_bus, TW), w.n_tiles), dim3(256), 0, stream, t, w);
  });
}

void launch_mismatch(const DevTopo& t, const DevWork& w, bool fp32, hipStream_t stream) {
  with_tile(w.tile, [&](auto tw) {
    constexpr int TW = decltype(tw)::value;
    dim3 grid(elem_blocks(t.n_bus, TW), w.n_tiles);
    if (fp32) hipLaunchKernelGGL((k_mismatch<float, TW>), grid, dim3(256), 0, stream, t, w);
    else hipLaunchKernelGGL((k_mismatch<double, TW>), grid, dim3(256), 0, stream, t, w);
  });
}

void launch_decide(const DevTopo& t, const DevWork& w, int iter, hipStream_t stream) {
  (void)t;
  const int threads = w.n_tiles * w.tile;
  hipLaunchKernelGGL(k_decide, dim3((threads + 255) / 256), dim3(256), 0, stream, w, iter);
}

void launch_fill(const DevTopo& t, const DevWork& w, bool fp32, hipStream_t stream) {
  with_tile(w.tile, [&](auto tw) {
    constexpr int TW = decltype(tw)::value;
    dim3 grid(elem_blocks(t.nnz_lu, TW), w.n_tiles);
    if (fp32) hipLaunchKernelGGL((k_fill<float, TW>), grid, dim3(256), 0, stream, t, w);
    else hipLaunchKernelGGL((k_fill<double, TW>), grid, dim3(256), 0, stream, t, w);
  });
}

size_t lu_shmem(const DevTopo& t, bool fp32, int tile) {
  return fp32 ? lu_bytes<float>(t, tile, true) : lu_bytes<double>(t, tile, true);
}

void launch_lu(const DevTopo& t, const DevWork& w, bool fp32, bool factor, int iter, hipStream_t stream) {
  with_tile(w.tile, [&](auto tw) {
    constexpr int TW = decltype(tw)::value;
    if (fp32) run_lu<float, TW>(t, w, factor, iter, stream);
    else run_lu<double, TW>(t, w, factor, iter, stream);
  });
}

void launch_refine(const DevTopo& t, const DevWork& w, double* dvr, double* dvi, hipStream_t stream) {
  with_tile(w.tile, [&](auto tw) {
    constexpr int TW = decltype(tw)::value;
    dim3 grid(elem_blocks(t.n_bus, TW), w.n_tiles);
    hipLaunchKernelGGL((k_dv<float, TW>), grid, dim3(256), 0, stream, t, w, dvr, dvi);
    hipLaunchKernelGGL((k_resid<float, TW>), grid, dim3(256), 0, stream, t, w, dvr, dvi);
  });
}

void launch_update(const DevTopo& t, const DevWork& w, bool fp32, bool mixed, hipStream_t stream) {
  with_tile(w.tile, [&](auto tw) {
    constexpr int TW = decltype(tw)::value;
    dim3 grid(elem_blocks(t.n_bus, TW), w.n_tiles);
    if (fp32) {
      if (mixed) hipLaunchKernelGGL((k_update<float, TW, true>), grid, dim3(256), 0, stream, t, w);
      else hipLaunchKernelGGL((k_update<float, TW, false>), grid, dim3(256), 0, stream, t, w);
    } else {
      hipLaunchKernelGGL((k_update<double, TW, false>), grid, dim3(256), 0, stream, t, w);
    }
  });
}

void launch_fallback(const DevTopo& t, const DevWork& w, hipStream_t stream) {
  with_tile(w.tile, [&](auto tw) {
    constexpr int TW = decltype(tw)::value;
    hipLaunchKernelGGL(k_fallback_restore<TW>, dim3(elem_blocks(t.n_bus, TW), w.n_tiles), dim3(256), 0, stream, t, w);
  });
  const int threads = w.n_tiles * w.tile;
  hipLaunchKernelGGL(k_fallback_reset, dim3((threads + 255) / 256), dim3(256), 0, stream, w);
}

void launch_finish(const DevWork& w, hipStream_t stream) {
  const int threads = w.n_tiles * w.tile;
  hipLaunchKernelGGL(k_finish, dim3((threads + 255) / 256), dim3(256), 0, stream, w);
}

void launch_island_max(const DevTopo& t, const DevWork& w, const int* comp, int n_comp, unsigned long long* out,
                       hipStream_t stream) {
  with_tile(w.tile, [&](auto tw) {
    constexpr int TW = decltype(tw)::value;
    hipLaunchKernelGGL(k_island_max<TW>, dim3(elem_blocks(t.n_bus, TW), w.n_tiles), dim3(256), 0, stream, t, w, comp,
                       n_comp, out);
  });
}

void launch_linear_guess(const DevTopo& t, const DevLg& lg, const DevTopo& lgt, const DevWork& w, void* lu,
                         void* y, int* lg_status, int* lg_iters, hipStream_t stream) {
  if (lg.m <= 0) return;
  with_tile(w.tile, [&](auto tw) {
    constexpr int TW = decltype(tw)::value;
    hipLaunchKernelGGL(k_lg_fill<TW>, dim3(elem_blocks(lg.nnz_lu, TW), w.n_tiles), dim3(256), 0, stream, t, lg, w,
                       reinterpret_cast<zd*>(lu), lg_status);
    hipLaunchKernelGGL(k_lg_rhs<TW>, dim3(elem_blocks(lg.m, TW), w.n_tiles), dim3(256), 0, stream, t, lg, w,
                       reinterpret_cast<zd*>(y));
    DevWork lw = w;
    lw.lu = lu;
    lw.y = y;
    lw.status = lg_status;
    lw.iters = lg_iters;
    run_lu<zd, TW>(lgt, lw, true, 0, stream);
    hipLaunchKernelGGL(k_lg_apply<TW>, dim3(elem_blocks(lg.m, TW), w.n_tiles), dim3(256), 0, stream, t, lg, w,
                       reinterpret_cast<const zd*>(y), lg_status);
  });
}

size_t lg_shmem(const DevTopo& lgt, int tile) { return lu_bytes<zd>(lgt, tile, true); }

void launch_lane0(const double* src, double* dst, int count, bool to_tile, int tile, hipStream_t stream) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_lane0, dim3((count + 255) / 256), dim3(256), 0, stream, src, dst, count, to_tile ? 1 : 0,
                     tile);
}

}  // namespace gx
