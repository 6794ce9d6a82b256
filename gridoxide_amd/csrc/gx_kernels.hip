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

#include <algorithm>
#include <climits>
#include <cmath>
#include <mutex>
#include <type_traits>

namespace gx {
namespace {

// Mixed precision: stagnation below this mismatch (pu) is attributed to the FP32 factors.
constexpr double kStallFloor = 1e-4;
// Mixed precision: a scenario refines its step (FP64 residual, FP32 correction solve) once its
// mismatch at that step is below this (pu); farther from the solution the FP32 solve's rounding is
// far below the error of the Newton step itself.  1e-2 rather than 1e-3: on ill-conditioned grids a
// warm-started scenario's first step (mismatch a few 1e-3) otherwise leaves a trajectory whose
// converged angles sit ~1e-8 rad from the FP64 ones (1e-2 and 1e-3 gave the same iteration counts on
// the 2,000-bus workload).
constexpr double kRefineBelow = 1e-2;

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
// (V scenarios per lane: the tile's TW V scenarios)
template <int TW, int V>
__device__ __forceinline__ bool tile_running_v(const int* status, int tile) {
  bool r = false;
#pragma unroll
  for (int e = 0; e < V; ++e) r = r || status[(tile * TW + sub<TW>()) * V + e] == kRunning;
  return __any(r);
}

// Whether scenario sc refines at this step (kRefineBelow): the step's mismatch is in the history.
__device__ __forceinline__ bool sc_refines(const DevWork& w, int sc, int iter) {
  const int it = w.iter_dev ? *w.iter_dev : iter;
  return w.status[sc] == kRunning && w.history[(size_t)sc * w.max_iter + it] < kRefineBelow;
}
// Whether any scenario of the tile refines (the tile then runs the sweep; every lane takes part).
// A scenario that does not refine ignores the sweep's results, so its own result does not depend
// on the scenarios it shares a tile with.
template <int TW>
__device__ __forceinline__ bool tile_refines(const DevWork& w, int tile, int iter) {
  return __any(sc_refines(w, tile * TW + sub<TW>(), iter));
}

// L+U values: the tiles go in blocks of B = lu_block(TW), each block holding its positions one
// after the other, each position the block's tiles side by side: entry p of lane s of tile t at
// ((t / B * nnz + p) * B + t % B) * TW + s.  A wavefront's gathers stay inside its block's region
// (nnz * B * TW values) instead of striding over all tiles' values, which keeps the address
// translations a wavefront needs few.  The allocation rounds the tile count up to a multiple of B.
__device__ __forceinline__ size_t lu_at(int p, int nnz, int tile, int TW) {
  const int B = lu_block(TW);
  return (((size_t)(tile / B) * nnz + p) * B + tile % B) * TW;
}
// Workgroup b of a one-dimensional grid of n runs on XCD b mod 8: number the tiles so that each XCD
// runs a contiguous range of them.
__device__ __forceinline__ int xcd_tile(int b, int n) {
  const int x = b & 7, i = b >> 3, q = n >> 3, r = n & 7;
  return x * q + (x < r ? x : r) + i;
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
// Two scenarios in one lane (pair tiles: a lane group of TW lanes holds 2 TW scenarios, a
// position's 2 TW values still contiguous), element by element.  Each load, store, LDS access and
// index computation of the LU kernel then serves two scenarios; a pair of pairs serves four (the
// refactor of quad tiles: one 16-byte access per lane).
template <typename S>
struct __attribute__((aligned(2 * sizeof(S)))) pair2 {
  S a, b;
};
using f2 = pair2<float>;
using d2 = pair2<double>;
using f4 = pair2<f2>;
template <typename S>
__device__ __forceinline__ pair2<S> operator-(pair2<S> x, pair2<S> y) { return {x.a - y.a, x.b - y.b}; }
template <typename S>
__device__ __forceinline__ pair2<S> operator+(pair2<S> x, pair2<S> y) { return {x.a + y.a, x.b + y.b}; }
template <typename S>
__device__ __forceinline__ pair2<S> operator*(pair2<S> x, pair2<S> y) { return {x.a * y.a, x.b * y.b}; }
template <typename S>
__device__ __forceinline__ pair2<S> operator/(pair2<S> x, pair2<S> y) { return {x.a / y.a, x.b / y.b}; }
// scenarios per lane of an LU element type
template <typename T>
struct lane_pack {
  static constexpr int V = 1;
};
template <typename S>
struct lane_pack<pair2<S>> {
  static constexpr int V = 2 * lane_pack<S>::V;
};
__device__ __forceinline__ bool is_zero(float v) { return v == 0.0f; }
__device__ __forceinline__ bool is_zero(double v) { return v == 0.0; }
__device__ __forceinline__ bool is_zero(zd v) { return v.re == 0.0 && v.im == 0.0; }
__device__ __forceinline__ bool finite_value(float v) { return isfinite(v); }
__device__ __forceinline__ bool finite_value(double v) { return isfinite(v); }
__device__ __forceinline__ bool finite_value(zd v) { return isfinite(v.re) && isfinite(v.im); }
// c - a b rounded once (a fused multiply-add), whatever the element type: the refactor's updates
// and the substitutions' multiply-subtracts give the same bits in every tile shape, pair tiles
// included, instead of depending on where the compiler contracts a product into a sum
__device__ __forceinline__ float fnma(float a, float b, float c) { return __builtin_fmaf(-a, b, c); }
__device__ __forceinline__ double fnma(double a, double b, double c) { return __builtin_fma(-a, b, c); }
__device__ __forceinline__ zd fnma(zd a, zd b, zd c) { return c - a * b; }
template <typename S>
__device__ __forceinline__ pair2<S> fnma(pair2<S> a, pair2<S> b, pair2<S> c) {
  return {fnma(a.a, b.a, c.a), fnma(a.b, b.b, c.b)};
}
// per-scenario flags of a lane's values (bit e: its scenario e): a zero pivot, a non-finite value
template <typename T>
__device__ __forceinline__ int zero_mask(T v) { return is_zero(v) ? 1 : 0; }
template <typename S>
__device__ __forceinline__ int zero_mask(pair2<S> v) { return zero_mask(v.a) | (zero_mask(v.b) << lane_pack<S>::V); }
template <typename T>
__device__ __forceinline__ int nonfinite_mask(T v) { return finite_value(v) ? 0 : 1; }
template <typename S>
__device__ __forceinline__ int nonfinite_mask(pair2<S> v) {
  return nonfinite_mask(v.a) | (nonfinite_mask(v.b) << lane_pack<S>::V);
}
// Keeps the compiler from moving memory accesses across this point.  Lanes of a wavefront that pass
// values through memory rely on the program order of the accesses; without it the compiler may, for
// example, sink a load into the branch that uses it, past another lane's store.
__device__ __forceinline__ void memory_order_point() { asm volatile("" ::: "memory"); }
// a real constant in every scenario of a value (zd: imaginary part zero)
template <typename T>
struct constant_of {
  static __device__ __forceinline__ T make(double x) { return T(x); }
};
template <>
struct constant_of<zd> {
  static __device__ __forceinline__ zd make(double x) { return zd{x, 0.0}; }
};
template <typename S>
struct constant_of<pair2<S>> {
  static __device__ __forceinline__ pair2<S> make(double x) { return {constant_of<S>::make(x), constant_of<S>::make(x)}; }
};
template <typename T>
__device__ __forceinline__ T zero_value() { return constant_of<T>::make(0.0); }
template <typename T>
__device__ __forceinline__ T minus_one() { return constant_of<T>::make(-1.0); }
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

// OR of a per-scenario flag mask over the lane groups (lanes with the same scenarios)
template <int TW>
__device__ __forceinline__ int any_groups(int v) {
  for (int d = TW; d < kLanes; d <<= 1) v |= __shfl_xor(v, d);
  return v;
}

template <int TW>
__device__ __forceinline__ bool all_groups(bool v) {
  int x = v ? 1 : 0;
  for (int d = TW; d < kLanes; d <<= 1) x &= __shfl_xor(x, d);
  return x != 0;
}

// The branch a scenario has out of service: its four Y-bus positions and stamps, subtracted from
// that scenario's Y values wherever a kernel reads Y (the pattern is unchanged).
struct Outage {
  int4 pos = make_int4(-1, -1, -1, -1);
  double re[4] = {0.0, 0.0, 0.0, 0.0}, im[4] = {0.0, 0.0, 0.0, 0.0};
};
__device__ __forceinline__ Outage load_outage(const DevTopo& t, const DevWork& w, int sc) {
  Outage o;
  if (w.outage) {
    const int e = w.outage[sc < w.n_scen ? sc : w.n_scen - 1];
    if (e >= 0) {
      o.pos = t.br_pos[e];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        o.re[q] = t.br_y[8 * e + 2 * q];
        o.im[q] = t.br_y[8 * e + 2 * q + 1];
      }
    }
  }
  return o;
}
// effective_injection (src/network.rs:600-611): the constant-power (P, Q) plus the bus's ZIP terms
// at its current voltage magnitude vm (constant power, S * vm, S * vm^2)
__device__ __forceinline__ void effective_injection(const DevTopo& t, int i, double vm, double& p, double& q) {
  if (t.zip) {
    const double* z = t.zip + 6 * (size_t)i;
    p = p + z[0] + z[2] * vm + z[4] * vm * vm;
    q = q + z[1] + z[3] * vm + z[5] * vm * vm;
  }
}
// Y value at position p for this scenario
__device__ __forceinline__ void y_at(const DevTopo& t, const Outage& o, int p, double& g, double& b) {
  g = t.y_g[p];
  b = t.y_b[p];
  if (p == o.pos.x) {
    g -= o.re[0];
    b -= o.im[0];
  } else if (p == o.pos.y) {
    g -= o.re[1];
    b -= o.im[1];
  } else if (p == o.pos.z) {
    g -= o.re[2];
    b -= o.im[2];
  } else if (p == o.pos.w) {
    g -= o.re[3];
    b -= o.im[3];
  }
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
__device__ __forceinline__ int elem_first(int bx = blockIdx.x) {
  return (bx * (blockDim.x / kLanes) + threadIdx.x / kLanes) * (kLanes / TW) + grp<TW>();
}
// The workgroup's (x, y) in a grid whose rows go whole to one XCD, so the rows' shared inputs (a
// tile's voltages, read for every neighbour of a bus) are fetched into that XCD's L2 once: the 8
// XCDs take workgroups round-robin in linear order, XCD c running linear ids c, c + 8, ..., and is
// given rows [c ny / 8, (c + 1) ny / 8), each row's columns in order.  A grid whose row count is
// not a multiple of 8 keeps its own order.
__device__ __forceinline__ void xcd_rows(int& x, int& y) {
  const int nx = gridDim.x, ny = gridDim.y;
  x = blockIdx.x;
  y = blockIdx.y;
  if (ny & 7) return;
  const int l = blockIdx.y * nx + blockIdx.x, c = l & 7, k = l >> 3;
  y = c * (ny >> 3) + k / nx;
  x = k % nx;
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
  if (blockIdx.x == 0)  // the mismatch history is defined (zero) past a scenario's last iteration
    for (int q = threadIdx.x; q < TW * w.max_iter; q += blockDim.x) w.history[(size_t)tile * TW * w.max_iter + q] = 0.0;
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

// The four Jacobian values of bus pair (i, k) from rectangular voltages (V_i, V_k, |V_k|) and
// Y(i, k) = g + jb, stored at their L+U positions pos = {H, N, M, L} (-1: none) with stride V; the
// diagonal pair adds the bus's calculated injection (pci, qci).  One definition for the fill kernel
// and the mismatch kernel that fills while it walks the Y rows, with no contraction of products into
// fused multiply-adds (which the compiler would otherwise choose per kernel), so both store the same
// bits.
template <typename T>
__device__ __forceinline__ void jac_store(T* __restrict__ lu, int V, int4 pos, double vri, double vii, double vrk,
                                          double vik, double vmk, double g, double b, bool diag, double pci,
                                          double qci) {
#pragma clang fp contract(off)
  const double cc = vri * vrk + vii * vik;
  const double ss = vii * vrk - vri * vik;
  const double hl = g * ss - b * cc;
  const double nm = g * cc + b * ss;
  // one FP64 division per pair (the mismatch kernel's Jacobian part is paced partly by its FP64
  // arithmetic); the products stay in FP64: in FP32 the cancellation in ss costs the FP32 Newton
  // steps enough accuracy that many more scenarios stall
  const double rk = 1.0 / vmk;
  double vh = hl, vn = nm * rk, vmm = -nm, vl = hl * rk;
  if (diag) {
    vh += -qci;
    vn += pci * rk;
    vmm += pci;
    vl += qci * rk;
  }
  if (pos.x >= 0) lu[(size_t)pos.x * V] = (T)vh;
  if (pos.y >= 0) lu[(size_t)pos.y * V] = (T)vn;
  if (pos.z >= 0) lu[(size_t)pos.z * V] = (T)vmm;
  if (pos.w >= 0) lu[(size_t)pos.w * V] = (T)vl;
}

// kFill: also the Jacobian values of every bus pair of the row into the L+U storage (what k_fill_tile
// writes, from the same Y values and voltages; the diagonal pair once the row's injection is known),
// so the row's voltages and Y values are read once for both.  The values of a tile whose scenarios
// all stop at this step are not used.
template <typename T, int TW, bool kFill = false>
__global__ __launch_bounds__(256) void k_mismatch(DevTopo t, DevWork w) {
  int bx, tile;
  xcd_rows(bx, tile);
  const int s = sub<TW>();
  if (!tile_running<TW>(w.status, tile)) return;
  const int sc = tile * TW + s;
  const size_t bb = at<TW>(tile, t.n_bus, 0) + s;
  const size_t ub = at<TW>(tile, t.n, 0) + s;
  T* __restrict__ y = reinterpret_cast<T*>(w.y);
  T* __restrict__ lu = kFill ? reinterpret_cast<T*>(w.lu) + lu_at(0, t.nnz_lu, tile, TW) + s : nullptr;
  constexpr int V = lu_block(TW) * TW;  // L+U values per position of a tile block
  const Outage og = load_outage(t, w, sc);
  double mx = 0.0;
  for (int i = elem_first<TW>(bx); i < t.n_bus; i += elem_stride<TW>()) {
    double ir = 0.0, ii = 0.0;
    const size_t o = bb + (size_t)i * TW;
    const double vri = w.vr[o], vii = w.vi[o];
    double gd = 0.0, bd = 0.0;
    int pd = -1;
    const int p2 = t.y_rowptr[i + 1];
    // the row's entries kMisBatch at a time: their columns, then their voltages and Y values, then
    // the sums in entry order (the loads of a batch in flight together)
    constexpr int kMisBatch = 4;
    for (int p0 = t.y_rowptr[i]; p0 < p2; p0 += kMisBatch) {
      int kk[kMisBatch];
#pragma unroll
      for (int u = 0; u < kMisBatch; ++u) kk[u] = t.y_col[p0 + u < p2 ? p0 + u : p0];
      double vrk[kMisBatch], vik[kMisBatch], gk[kMisBatch], bk[kMisBatch], vmk[kMisBatch];
#pragma unroll
      for (int u = 0; u < kMisBatch; ++u) {
        const int p = p0 + u < p2 ? p0 + u : p0;
        const size_t oj = bb + (size_t)kk[u] * TW;
        y_at(t, og, p, gk[u], bk[u]);
        vrk[u] = w.vr[oj];
        vik[u] = w.vi[oj];
        if constexpr (kFill) vmk[u] = w.vm[oj];
      }
#pragma unroll
      for (int u = 0; u < kMisBatch; ++u) {
        if (p0 + u >= p2) break;
        const int p = p0 + u, k = kk[u];
        const double g = gk[u], b = bk[u], vr = vrk[u], vi = vik[u];
        ir += g * vr - b * vi;
        ii += g * vi + b * vr;
        if constexpr (kFill) {
          if (k == i) {
            gd = g;
            bd = b;
            pd = p;
          } else {
            jac_store(lu, V, t.y_lupos[p], vri, vii, vr, vi, vmk[u], g, b, false, 0.0, 0.0);
          }
        }
      }
    }
    const double vr = vri, vi = vii;
    const double pcal = vr * ir + vi * ii;
    const double qcal = vi * ir - vr * ii;
    w.pc[o] = pcal;
    w.qc[o] = qcal;
    if constexpr (kFill) {
      if (pd >= 0) jac_store(lu, V, t.y_lupos[pd], vri, vii, vri, vii, w.vm[o], gd, bd, true, pcal, qcal);
    }
    double pe = w.ps[o], qe = w.qs[o];
    effective_injection(t, i, w.vm[o], pe, qe);
    const int rp = t.rpos_p[i];
    if (rp >= 0) {
      const double fp = pe - pcal;
      w.f[ub + (size_t)rp * TW] = fp;
      y[ub + (size_t)rp * TW] = (T)fp;
      mx = fmax(mx, fabs(fp));
    }
    const int rq = t.rpos_q[i];
    if (rq >= 0) {
      const double fq = qe - qcal;
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
  if (w.iter_dev) iter = *w.iter_dev;
  const int sc = blockIdx.x * blockDim.x + threadIdx.x;
  if (sc >= w.n_tiles * w.tile) return;
  if (w.status[sc] != kRunning) return;
  const double m = __longlong_as_double((long long)w.maxmis[sc]);
  w.maxmis[sc] = 0ull;
  w.history[(size_t)sc * w.max_iter + iter] = m;
  if (m < w.tol) {
    w.status[sc] = kConverged;
    w.iters[sc] = iter + 1;
  } else if (w.sing0 && w.sing0[sc]) {
    // its first Jacobian had no valid pivot order: Singular at the first solve, as when
    // PersistentSolver's first factorization fails (src/solver.rs:711-716)
    w.status[sc] = kSingular;
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
  if (sc == 0) {
    for (int it = 0; it < w.max_iter; ++it) w.running[it] = 0;
    if (w.iter_dev) *w.iter_dev = 0;
  }
  if (sc >= w.n_tiles * w.tile) return;
  const int st = w.status[sc];
  if (st != kStalled && st != kSingular) return;
  if (w.sing0 && w.sing0[sc]) return;  // (Singular in either precision)
  w.status[sc] = kRunning;
  w.iters[sc] = 0;
  w.maxmis[sc] = 0ull;
  for (int it = 0; it < w.max_iter; ++it) w.history[(size_t)sc * w.max_iter + it] = 0.0;
  atomicAdd(w.fallback, 1);
}

// Jacobian values into the packed L+U storage, in a branch-free form of the eight formulas: with
// C = V_i V_k cos(t_i - t_k) and S = V_i V_k sin(t_i - t_k) (C = V_i^2, S = 0 on the diagonal),
//   H = G S - B C,  N = (G C + B S) / V_k,  M = -(G C + B S),  L = (G S - B C) / V_k,
// plus the injection term of the diagonal kinds: H -= Q_i, N += P_i / V_i, M += P_i,
// L += Q_i / V_i (src/jacobian.rs:220-243).
template <typename T, int TW>
__global__ __launch_bounds__(256) void k_fill(DevTopo t, DevWork w) {
  // one L+U position per workgroup, one tile per lane group: the position's kind, buses and
  // admittance are uniform and the stores of the workgroup's tiles are contiguous
  const int p = blockIdx.x, tile = blockIdx.y * (256 / TW) + (int)threadIdx.x / TW, s = sub<TW>();
  if (tile >= w.n_tiles) return;
  const int sc = tile * TW + s;
  bool run = false;  // any scenario of the tile still iterating (the LU kernel skips the others)
  for (int q = 0; q < TW; ++q) run = run || w.status[tile * TW + q] == kRunning;
  if (!run) return;
  T* __restrict__ lu = reinterpret_cast<T*>(w.lu) + lu_at(p, t.nnz_lu, tile, TW) + s;
  const int kind = t.lu_kind[p];
  double v = 0.0;
  if (kind != kFill) {
    const size_t bb = at<TW>(tile, t.n_bus, 0) + s;
    const Outage og = load_outage(t, w, sc);
    const int i = t.lu_bi[p], k = t.lu_bk[p];
    double g = t.lu_g[p], b = t.lu_b[p];
    if (w.outage) y_at(t, og, t.lu_yp[p], g, b);
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
  *lu = (T)v;
}

// Jacobian fill by bus (JacobianPattern::fill_into, src/jacobian.rs:210-250, in rectangular form),
// block of tiles by block: the L+U values of a block's B tiles lie side by side, B * TW values per
// position, so a lane group of B * TW lanes (the block's scenarios) writes whole runs; a wavefront
// takes one block and 64 / (B * TW) consecutive row buses (one per lane group).  V, P and Q of a row
// bus are read once and each neighbour k's V once for its up to four entries H, N, M, L (i, k); the
// neighbours of a bus sit a few buses away in the same tile's arrays.
template <typename T, int TW>
__global__ __launch_bounds__(256) void k_fill_tile(DevTopo t, DevWork w) {
  // a lane group spans a whole block when it fits a wavefront, else one tile (the wide tile)
  constexpr int B = lu_block(TW), V = B * TW, LW = V <= kLanes ? V : TW, GB = kLanes / LW;
  const int lane = lane_id(), grp_b = lane / LW, sub_b = lane % LW;
  int bx, by;
  xcd_rows(bx, by);
  // (the last block may hold tiles past n_tiles: they read the last tile's inputs, their values are
  // never used)
  int tile = V <= kLanes ? by * B + sub_b / TW : by;
  tile = tile < w.n_tiles ? tile : w.n_tiles - 1;
  const int s = sub_b % TW;
  bool run = false;  // any scenario of the group's tiles still iterating (the LU kernel skips the others)
  for (int q = 0; q < LW; ++q) {
    const int sq = by * LW + q;
    run = run || (sq < w.n_tiles * TW && w.status[sq] == kRunning);
  }
  if (!run) return;
  const int sc = tile * TW + s;
  const size_t bb = at<TW>(tile, t.n_bus, 0) + s;
  const Outage og = load_outage(t, w, sc);
  T* __restrict__ lu = reinterpret_cast<T*>(w.lu) + (V <= kLanes ? (size_t)by * t.nnz_lu * V + sub_b
                                                                  : lu_at(0, t.nnz_lu, tile, TW) + s);
  const int first = (bx * (blockDim.x / kLanes) + (int)threadIdx.x / kLanes) * GB + grp_b;
  const int stride = gridDim.x * (blockDim.x / kLanes) * GB;
  for (int i = first; i < t.n_bus; i += stride) {
    const int e0 = t.fb_ptr[i], e1 = t.fb_ptr[i + 1];
    const size_t oi = bb + (size_t)i * TW;
    const double vri = w.vr[oi], vii = w.vi[oi], pci = w.pc[oi], qci = w.qc[oi];
    for (int e = e0; e < e1; ++e) {
      const int k = t.fb_k[e];
      const int4 pos = t.fb_pos[e];
      double g = t.fb_g[e], b = t.fb_b[e];
      if (w.outage) y_at(t, og, t.fb_yp[e], g, b);
      const size_t ok = bb + (size_t)k * TW;
      jac_store(lu, V, pos, vri, vii, w.vr[ok], w.vi[ok], w.vm[ok], g, b, k == i, pci, qci);
    }
  }
}

// Zero into the positions that are no Jacobian entry (structural fill), ahead of each refactor.
template <typename T, int TW>
__global__ __launch_bounds__(256) void k_fill_zero(DevTopo t, DevWork w) {
  // one block of B tiles per workgroup row (grid.y); a thread zeroes one fill position's B * TW
  // contiguous values, so a wavefront's stores cover neighbouring positions
  constexpr int B = lu_block(TW), V = B * TW;
  const int blk = blockIdx.y;
  const int lane0 = blk * V + (int)threadIdx.x;  // (the block's scenarios: threads 0..V-1)
  const bool mine = threadIdx.x < V && lane0 < w.n_tiles * TW && w.status[lane0] == kRunning;
  if (!__syncthreads_or(mine)) return;
  T* __restrict__ base = reinterpret_cast<T*>(w.lu) + (size_t)blk * t.nnz_lu * V;
  for (int f = blockIdx.x * 256 + (int)threadIdx.x; f < t.n_fz; f += gridDim.x * 256) {
    T* __restrict__ v = base + (size_t)t.fz_pos[f] * V;
#pragma unroll
    for (int i = 0; i < V; ++i) v[i] = (T)0;
  }
}

// Entries of a column a lane holds in registers (longer columns take a slower loop).
template <int G>
constexpr int stage_regs() {
  return 32 / G + 1;
}

__device__ __forceinline__ int rl(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

// c ? a : b on values (a conditional between two lvalues of a struct type can become a select of
// their addresses, which keeps both in scratch memory)
template <typename T>
__device__ __forceinline__ T pick(bool c, T a, T b) { return c ? a : b; }

// f(std::integral_constant<int, k>{}) for k = 0 .. N-1, expanded at compile time
template <typename F, int... K>
__device__ __forceinline__ void static_for_seq(F& f, std::integer_sequence<int, K...>) {
  (f(std::integral_constant<int, K>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_seq(f, std::make_integer_sequence<int, N>{});
}

// Column metadata through the vector memory path (every lane reads the same 16 bytes, a few
// columns ahead): waits on it are counted with the other vector loads.  A scalar-cache load would
// share its wait counter with the LDS operations, so every LDS wait of the column would also wait
// for it.  The empty asm only moves the index into a vector register, which keeps the load a
// vector load; uni4 takes the (uniform) value back into scalar registers where it is used.
// A load the compiler is told is global memory (a generic pointer may otherwise become a flat
// access, which also counts as an LDS access in the wait counters).
__device__ __forceinline__ int global_int(const int* p) {
  return *(const __attribute__((address_space(1))) int*)(p);
}
__device__ __forceinline__ int4 vload4(const int4* p, int i) {
  int j = i;
  asm("" : "+v"(j));
  return p[j];
}
__device__ __forceinline__ int4 uni4(int4 v) {
  return make_int4(__builtin_amdgcn_readfirstlane(v.x), __builtin_amdgcn_readfirstlane(v.y),
                   __builtin_amdgcn_readfirstlane(v.z), __builtin_amdgcn_readfirstlane(v.w));
}

// Waits for every outstanding memory operation: closes the rare variable-length paths (long
// columns, overflow gathers) so that the common path's waits stay exact.
__device__ __forceinline__ void drain() { __builtin_amdgcn_s_waitcnt(0); }

// A triangular solve by steps (item encoding at kBackItems) for one tile: every lane group runs its
// pieces of rows one after another as a dependent chain in registers.  kBackAhead register
// sets take turns: after a set's step has run, it loads the values of the step kBackAhead later
// (their descriptors were loaded kBackAhead steps earlier still), so no load is waited for soon after
// it was issued and no in-flight load is copied between registers.  x of rows that ended recently
// comes from the LDS ring, older x from y; partial sums of row pieces only pass through the ring.
// kDivide: a row's END divides by the pivot (back substitution); else END keeps the sum (forward
// substitution with the unit L).
//   items: this lane group's first item; lub/pstb: the lane's L+U base and position stride (bytes);
//   yl: the lane's y (offsets row * TW); rl: the lane's ring (offsets slot * S); R ring rows;
//   bad: the lane's scenarios with a non-finite solution are flagged (nonfinite_mask)
template <typename T, int TW, bool kDivide>
__device__ __forceinline__ void level_solve(const int2* __restrict__ bib, int g, int nst, int R, int nl,
                                            const char* __restrict__ lbase, unsigned pstb, char* __restrict__ ybase,
                                            unsigned loff, T* rl, int& bad) {
  constexpr int G = kLanes / TW;
  constexpr int S = TW + 1;
  constexpr int P = kBackItems, D = kBackAhead;
  constexpr unsigned kPayload = (1u << 29) - 1;
  const T zero = zero_value<T>();
  // uniform bases plus 32-bit lane offsets (loff: the lane's byte offset within a position and
  // within a row of y), so the gathers can take the scalar-base addressing form
  auto LU = [&](unsigned p) -> const T& { return *reinterpret_cast<const T*>(lbase + (p * pstb + loff)); };
  auto Y = [&](unsigned e) -> T& { return *reinterpret_cast<T*>(ybase + (e * (unsigned)sizeof(T) + loff)); };
  const unsigned pos_max = (unsigned)(nl - 1);
  // a set holds the values of its step and the descriptors of that step and of the step D later in
  // da / db, whose roles alternate from one turn of the set to the next (so that the loop needs no
  // register copies, which would wait for the loads in flight)
  struct BSet {
    int2 da[P], db[P];
    T u[P], y[P];
  };
  auto dsc_issue = [&](int q, int2 (&d)[P]) __attribute__((always_inline)) {
    const int qq = q < nst ? q : nst - 1;
#pragma unroll
    for (int i = 0; i < P; ++i)
      d[i] = *reinterpret_cast<const int2*>(reinterpret_cast<const char*>(bib) + (unsigned)((qq * P + i) * G + g) * 8u);
  };
  auto val_issue = [&](const int2 (&d)[P], BSet& b) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < P; ++i) {
      // (a START's payload is an LDS offset: clamped into the L+U range, the value is unused)
      const unsigned p = (unsigned)d[i].x & kPayload;
      b.u[i] = LU(p < pos_max ? p : pos_max);
      b.y[i] = Y((unsigned)(d[i].y >= 0 ? d[i].y : 0));
    }
  };
  T acc = zero;
  unsigned xslot = 0, xrow = 0;  // ring and y offsets of the row being run (from its START)
  rl[(unsigned)((R - 1) * S)] = zero;  // the zero row (END and PAD items multiply by it)
  // step q with the set's values and descriptors cur; then the values of step q + D (descriptors
  // nxt) and the descriptors of step q + 2D into cur
  auto run = [&](int q, BSet& b, int2 (&cur)[P], int2 (&nxt)[P]) __attribute__((always_inline)) {
    T lx[P];
#pragma unroll
    for (int i = 0; i < P; ++i) lx[i] = rl[(unsigned)(cur[i].y < 0 ? cur[i].y & kPayload : 0)];
    memory_order_point();  // every LDS read of the step before any of its writes
    T pend = zero, pend_u = zero;
    unsigned pend_y = 0, pend_s = 0;
    bool ended = false;
#pragma unroll
    for (int i = 0; i < P; ++i) {
      const unsigned a = (unsigned)cur[i].x;
      const int bb = cur[i].y;
      const unsigned kind = a >> 29;
      const bool start = kind == (unsigned)kBackStart;
      xslot = start ? (a & kPayload) : xslot;
      xrow = start ? (unsigned)bb : xrow;
      // ENTRY: acc - U(k, c) x_c; END and PAD read the zero row (acc unchanged); START: y_k or zero;
      // COMBINE: acc + a partial
      const T xv = pick(bb < 0, lx[i], b.y[i]);
      const T v = fnma(b.u[i], xv, acc);
      const T w = acc + xv;
      acc = pick(start, xv, pick(kind == (unsigned)kBackCombine, w, v));
      const bool end = kind == (unsigned)kBackEnd;  // (at most one per lane group and step)
      pend = pick(end, acc, pend);
      pend_u = pick(end, b.u[i], pend_u);
      pend_y = end ? xrow : pend_y;
      pend_s = end ? xslot : pend_s;
      ended = ended || end;
    }
    if (ended) {  // the step's piece end: a partial into the ring; x_k into y and the ring
      const bool partial = (int)pend_y < 0;  // (its START named the zero row)
      const T xk = kDivide && !partial ? pend / pend_u : pend;
      bad |= nonfinite_mask(xk);
      if (!partial) Y(pend_y) = xk;
      rl[pend_s] = xk;
    }
    memory_order_point();  // the step's stores before the next loads
    val_issue(nxt, b);
    dsc_issue(q + 2 * D, cur);
  };
  // (every index into sets is a compile-time constant, so the sets live in registers whatever the
  // size of the unrolled loop body)
  BSet sets[D];
  static_for<D>([&](auto k) { dsc_issue(k, sets[k].da); });
  static_for<D>([&](auto k) { dsc_issue(D + k, sets[k].db); });
  static_for<D>([&](auto k) { val_issue(sets[k].da, sets[k]); });
  drain();  // (the loop's waits are counted from what the loop itself issues)
  for (int q = 0; q < nst; q += 2 * D) {
    static_for<D>([&](auto k) { run(q + k, sets[k], sets[k].da, sets[k].db); });
    static_for<D>([&](auto k) { run(q + D + k, sets[k], sets[k].db, sets[k].da); });
  }
}

// Refactor (optional) + forward/back substitution; one wavefront per tile.
//   LDS: two staged columns xc[2][cap+1][TW+1] (column k in buffer k & 1; the previous column stays
//   readable for the updates that use it) and ring[R][TW+1] (the right-hand side rows inside the
//   moving window; row r lives in slot r mod R, R = W + kRingGap + 1 a power of two).
// Refactor of column k: stage A(:,k), run its update rounds (host-scheduled, one item per lane
// group and round: x_t -= L(i,j) * x_u), divide the L part by the pivot, write back.
// Forward: at column k the ring holds rows [k - B, k + W] (B = kRingGap); y_k = ring[k]; L(:,k)
// pushes into the ring rows i <= k + W; row k + W + 1 then enters the ring (taking the slot of
// row k - B) as f minus the contributions of its L entries farther than W, gathered at that
// moment.  Back substitution mirrors this with U: the ring holds rows [k - W, k + B].
// Memory: column metadata is loaded two columns ahead.  The update rounds
// form one stream of chunks over all columns: each chunk step issues the slots of the chunk two
// ahead and the L values of the next chunk (whose slots arrived during the previous step), so no
// chunk waits for its own loads, in long columns as across columns.  A column's staged values and
// rows, its far entries and its entering row's f are issued one column ahead.  Addresses are a
// uniform base plus a 32-bit lane offset.  Every load on the common path is unconditional (clamped
// or padded addresses), so the compiler's waits are exact; the rare long-column tails load in
// fixed batches.  L values of column k-1 (not yet written back when the loads of column k's first
// chunk are issued) are read from LDS instead.
template <typename T, int TW, bool kFactor>
__device__ __forceinline__ void lu_tile(const DevTopo& t, const DevWork& w, int tile, int cap, int iter) {
  constexpr int G = kLanes / TW;
  constexpr int S = TW + 1;
  constexpr int B = kRingGap;
  constexpr int NB = kChunkRounds;
  constexpr int MS = stage_regs<G>();
  constexpr int MT = 4;  // long-column tail entries per lane group and batch
  const int W = t.window;
  const int R = W + B + 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* xcb = reinterpret_cast<T*>(smem_raw);
  const int xstride = (cap + 1) * S;
  T* ring = xcb + (kFactor ? 2 * (size_t)xstride : 0);
  const int g = grp<TW>(), s = sub<TW>();
  if (!tile_running<TW>(w.status, tile)) return;
  if (!kFactor && !w.swept[tile]) return;  // (the mixed correction solve)
  const int sc = tile * TW + s;
  const int n = t.n;
  const int nl = t.nnz_lu;
  // this tile's arrays (uniform bases); entry e of lane s at e * TW + s
  char* __restrict__ lub = reinterpret_cast<char*>(reinterpret_cast<T*>(w.lu) + lu_at(0, t.nnz_lu, tile, TW) + s);
  const char* __restrict__ lbase = reinterpret_cast<const char*>(reinterpret_cast<T*>(w.lu) + lu_at(0, t.nnz_lu, tile, TW));
  T* __restrict__ yb = reinterpret_cast<T*>(w.y) + (size_t)tile * n * TW;
  const unsigned pstb = (unsigned)(lu_block(TW) * TW * sizeof(T));  // bytes between consecutive positions
  // (one multiply-add per address: position x stride + this lane's base)
  auto LU = [&](int p) -> T& { return *reinterpret_cast<T*>(lub + (size_t)(unsigned)p * pstb); };
  auto Y = [&](int r) -> T& { return yb[(unsigned)(r * TW + s)]; };
  const int* __restrict__ rowp = t.lu_row;
  bool ok = true;        // per lane: no zero pivot, finite solution
  bool overflow = false; // a column longer than the staging area (the same for every lane)
  auto slot = [&](int r) { return (r & (R - 1)) * S + s; };
  auto cl = [](int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); };
  const T zero = zero_value<T>();
  // Lane-group strided loop body(i) over i = base + off + G*j < end, with the loop control on
  // the uniform base (0 <= off < G) and the lane part predicated: every lane leaves the loop
  // together, so the loop exit never runs with an empty exec mask (register copies that the
  // compiler places in such an exit block would be skipped).
  auto strided = [&](int base, int off, int end, auto&& body) {
    for (int b = base; b < end; b += G) {
      const int i = b + off;
      if (i < end) body(i);
    }
  };

  // ------------------------------------------------ forward pass (+ refactor)
  strided(0, g, W + 1 < n ? W + 1 : n, [&](int r) { ring[slot(r)] = Y(r); });
  // column metadata a = {c0, len, ul, chunk0}, f = {rounds, overflow range, far flag}: m*1 = next
  // column, m*2 = the one after; fr1/fr2 = f of the rows entering after those columns
  int4 ma1 = vload4(t.col_a, 0), ma2 = vload4(t.col_a, cl(1, n - 1));
  int4 mf1 = vload4(t.col_f, 0), mf2 = vload4(t.col_f, cl(1, n - 1));
  T fr1 = Y(cl(W + 1, n - 1)), fr2 = Y(cl(W + 2, n - 1));

  // staged entries: factor -> entries o = first_of_group(c0, g) + G*i of the whole column (owned
  // positions); solve only -> L entries o = ul + 1 + g + G*i
  T sv[MS], cv[MS];
  int sr[MS], cr[MS];
  auto stage_issue = [&](const int4& m) {
    const int c0 = m.x, len = m.y, ul = m.z;
    const int f0 = kFactor ? first_of_group<G>(c0, g) : ul + 1 + g;
#pragma unroll
    for (int i = 0; i < MS; ++i) {
      const int o = f0 + G * i;
      const int p = c0 + (o < len ? o : 0);
      sv[i] = LU(p);
      sr[i] = rowp[p];
    }
  };
  // update-round chunk stream: sA/vA = the chunk being run (slots, L values), sB = the next one
  // (its values vB are issued while sA runs), sC = the one after
  int2 sA[NB], sB[NB], sC[NB];
  T vA[NB], vB[NB];
  const int2* __restrict__ slb = t.slots + g;
  auto slots_issue = [&](int q, int2 (&d)[NB]) {
#pragma unroll
    for (int r = 0; r < NB; ++r) d[r] = slb[(unsigned)((q * NB + r) * G)];
  };
  auto vals_issue = [&](const int2 (&sm)[NB], T (&d)[NB]) {
#pragma unroll
    for (int r = 0; r < NB; ++r) d[r] = LU(sm[r].x);
  };
  // far entries of the row entering after a column: f0s/f0l/f0y = column k (slot, L value, y value),
  // f1s/f1l/f1y = column k+1, f2s = slot of column k+2
  int2 f0s, f1s, f2s;
  T f0l, f0y, f1l, f1y;
  auto fslot_issue = [&](int kk) { return t.fslot[(unsigned)(cl(kk, n - 1) * G + g)]; };
  auto fval_issue = [&](int2 fs, T& lv, T& yv) {
    lv = LU(fs.x < 0 ? 0 : fs.x);
    yv = Y(fs.y < 0 ? 0 : fs.y);
  };
#pragma unroll
  for (int i = 0; i < MS; ++i) cv[i] = zero;
  stage_issue(uni4(ma1));
  if (kFactor) {
    slots_issue(0, sA);
    slots_issue(1, sB);
    vals_issue(sA, vA);
  }
  f1s = fslot_issue(0);
  f2s = fslot_issue(1);
  fval_issue(f1s, f1l, f1y);
  // (the loop's waits are counted from what the previous column issued; without this wait the
  // loads above, which the first column reads, would make every column's waits stricter)
  drain();
  int prev_c0 = 0;  // first packed position of column k-1
  for (int k = 0; k < n; ++k) {
    const int4 mk = uni4(ma1), fk = uni4(mf1);
    ma1 = ma2;
    mf1 = mf2;
    ma2 = vload4(t.col_a, cl(k + 2, n - 1));
    mf2 = vload4(t.col_f, cl(k + 2, n - 1));
    const int c0 = mk.x, len = mk.y, ul = mk.z, ch0 = mk.w;
    const int nr = fk.x, oq0 = fk.y, oq1 = fk.z, far = fk.w;
    // staged entries o = fb + fo + G*i (fb uniform, 0 <= fo < G); the rest from fb + G*MS on
    const int fb = kFactor ? 0 : ul + 1, fo = kFactor ? first_of_group<G>(c0, g) : g;
    const int f0 = fb + fo;
    const int tail = fb + G * MS;  // first uniform base past the staged entries
    T* xc = xcb + (k & 1) * xstride;
    const T* xprev = xcb + ((k & 1) ^ 1) * xstride;
    // this column's staged data (the registers are refilled for the next column below)
#pragma unroll
    for (int i = 0; i < MS; ++i) cr[i] = sr[i];
    if (kFactor) {
      if (len > cap) {  // the host sizes cap to the longest column; kept as a guard
        overflow = true;
        break;
      }
#pragma unroll
      for (int i = 0; i < MS; ++i) {
        const int o = f0 + G * i;
        if (o < len) xc[o * S + s] = sv[i];
      }
      for (int b = tail; b < len; b += G * MT) {  // long column: the rest, in batches
        T tv[MT];
#pragma unroll
        for (int u = 0; u < MT; ++u) {
          const int o = b + G * u + fo;
          tv[u] = LU(c0 + (o < len ? o : 0));
        }
#pragma unroll
        for (int u = 0; u < MT; ++u) {
          const int o = b + G * u + fo;
          if (o < len) xc[o * S + s] = tv[u];
        }
      }
    } else {
#pragma unroll
      for (int i = 0; i < MS; ++i) cv[i] = sv[i];
    }
    // loads for the next columns
    f0s = f1s;
    f0l = f1l;
    f0y = f1y;
    f1s = f2s;
    const T cfr = fr1;
    fr1 = fr2;
    fr2 = Y(cl(k + W + 3, n - 1));
    stage_issue(uni4(ma1));
    fval_issue(f1s, f1l, f1y);
    f2s = fslot_issue(k + 2);
    if (kFactor) {
      const int nch = nr > NB ? (nr + NB - 1) / NB : 1;
      for (int c = 0; c < nch; ++c) {
        slots_issue(ch0 + c + 2, sC);
        vals_issue(sB, vB);
        const int rleft = nr - c * NB;
#pragma unroll
        for (int r = 0; r < NB; ++r) {
          if (r < rleft) {
            const int wd = sA[r].y;
            const int tg = wd & 0xffff, su = (wd >> 16) & 0x7fff;
            // (conditional statements, not conditional lvalues: a select between addresses would
            // move the register arrays into private memory)
            T lv = vA[r];
            if (wd < 0) lv = xprev[sA[r].x * S + s];
            const T xt = xc[tg * S + s], xu = xc[su * S + s];
            xc[tg * S + s] = xt - lv * xu;
          }
        }
        if (c + 1 < nch) {  // next chunk of this column (the last one rotates after the column)
#pragma unroll
          for (int r = 0; r < NB; ++r) {
            sA[r] = sB[r];
            sB[r] = sC[r];
            vA[r] = vB[r];
          }
        }
      }
      const T ukk = xc[ul * S + s];
      if (is_zero(ukk)) ok = false;
#pragma unroll
      for (int i = 0; i < MS; ++i) {
        const int o = f0 + G * i;
        if (o < len) {
          T v = xc[o * S + s];
          if (o > ul) {
            v = v / ukk;
            xc[o * S + s] = v;
          }
          cv[i] = v;
        }
      }
      // write back (unused slots rewrite the diagonal: the same value its owner stores)
#pragma unroll
      for (int i = 0; i < MS; ++i) {
        const int o = f0 + G * i;
        const int oo = o < len ? o : ul;
        LU(c0 + oo) = o < len ? cv[i] : ukk;
      }
      for (int b = tail; b < len; b += G * MT) {
        T tv[MT];
#pragma unroll
        for (int u = 0; u < MT; ++u) {
          const int o = b + G * u + fo;
          tv[u] = xc[(o < len ? o : ul) * S + s];
        }
#pragma unroll
        for (int u = 0; u < MT; ++u) {
          const int o = b + G * u + fo;
          if (o < len) {
            T v = tv[u];
            if (o > ul) {
              v = v / ukk;
              xc[o * S + s] = v;
            }
            LU(c0 + o) = v;
          }
        }
      }
    }
    // forward substitution: y_k is final (every lane group stores the same value)
    const T yk = ring[slot(k)];
    Y(k) = yk;
    // the row entering the window after column k: f minus its far L entries
    const int rin = k + W + 1;
    if (rin < n) {
      T enter = cfr;
      if (far) {
        T acc = zero;
        if (f0s.x >= 0) {
          const int j = f0s.y;
          T lv = f0l;
          if (kFactor && j == k) lv = xc[(f0s.x - c0) * S + s];
          if (kFactor && j == k - 1) lv = xprev[(f0s.x - prev_c0) * S + s];
          T yj = f0y;
          if (j >= k - B) yj = ring[slot(j)];
          acc = lv * yj;
        }
        strided(oq0, g, oq1, [&](int q) {  // entries beyond the first G
          const int p = t.fo_pos[q], j = t.fo_col[q];
          T yj, lv;
          if (j > k - B) yj = ring[slot(j)];
          else yj = Y(j);
          if (kFactor && j == k) lv = xc[(p - c0) * S + s];
          else lv = LU(p);
          acc = acc + lv * yj;
        });
        enter = cfr - group_sum<TW>(acc);
      }
      if (g == k % G) ring[slot(rin)] = enter;
    }
    // push L(:,k) into the window rows (rows farther than W are gathered when they enter); the
    // rows of one column are distinct, so all reads go first
    {
      T rv[MS];
#pragma unroll
      for (int i = 0; i < MS; ++i) rv[i] = ring[slot(cr[i])];
#pragma unroll
      for (int i = 0; i < MS; ++i) {
        const int o = f0 + G * i;
        if (o > ul && o < len && cr[i] - k <= W) ring[slot(cr[i])] = rv[i] - cv[i] * yk;
      }
    }
    for (int b = tail; b < len; b += G * MT) {
      int ri[MT];
      T lv[MT], rv[MT];
#pragma unroll
      for (int u = 0; u < MT; ++u) {
        const int o = b + G * u + fo;
        const int oc = o < len ? o : ul;
        ri[u] = rowp[c0 + oc];
        lv[u] = kFactor ? xc[oc * S + s] : LU(c0 + oc);
      }
#pragma unroll
      for (int u = 0; u < MT; ++u) rv[u] = ring[slot(ri[u])];
#pragma unroll
      for (int u = 0; u < MT; ++u) {
        const int o = b + G * u + fo;
        if (o > ul && o < len && ri[u] - k <= W) ring[slot(ri[u])] = rv[u] - lv[u] * yk;
      }
    }
    if (kFactor) {  // the chunk stream moves on to the next column's first chunk
#pragma unroll
      for (int r = 0; r < NB; ++r) {
        sA[r] = sB[r];
        sB[r] = sC[r];
        vA[r] = vB[r];
      }
    }
    prev_c0 = c0;
    if ((k + 1) % B == 0) workgroup_fence();
  }
  workgroup_fence();

  // ------------------------------------------------ back substitution
  if (!overflow && t.back_items) {
    int bad = 0;
    level_solve<T, TW, true>(t.back_items, g, t.back_steps, R, nl, lbase, pstb, reinterpret_cast<char*>(yb),
                             (unsigned)(s * sizeof(T)), ring + s, bad);
    if (bad) ok = false;
  } else if (!overflow) {
    strided(n - 1 - W > 0 ? n - 1 - W : 0, g, n, [&](int r) { ring[slot(r)] = Y(r); });
    // column metadata three columns ahead; the U entries, the pivot, the entering
    // row's y and its first far entries two columns ahead, in two register sets that alternate
    // between columns (the loop is unrolled by two, so a set is refilled only after it has been
    // consumed and no in-flight load is copied between registers)
    int4 ba1 = vload4(t.col_a, n - 1), ba2 = vload4(t.col_a, cl(n - 2, n - 1));
    int4 ba3 = vload4(t.col_a, cl(n - 3, n - 1));
    int4 bb1 = vload4(t.col_b, n - 1), bb2 = vload4(t.col_b, cl(n - 2, n - 1));
    struct USet {
      T uv[MS];
      int ur[MS];
      T pv, yr, fl, fx;  // pivot, entering row's y, first far U value and its x (or y) value
      int2 fs;           // first far entry of this lane group {position, column}
    };
    USet A, Bs;
    // (the far-entry slot of a set is loaded one step before the set, so that its value loads do
    // not wait for it)
    auto bslot_of = [&](int kk) { return t.bslot[(unsigned)(cl(kk, n - 1) * G + g)]; };
    auto u_issue = [&](int kk, const int4& m, USet& d, int2 fs) {
      const int c0 = m.x, ul = m.z;
#pragma unroll
      for (int i = 0; i < MS; ++i) {
        const int u = g + G * i;
        const int p = c0 + (u < ul ? u : 0);
        d.uv[i] = LU(p);
        d.ur[i] = rowp[p];
      }
      d.pv = LU(c0 + ul);
      d.yr = Y(cl(kk - W - 1, n - 1));
      d.fs = fs;
      d.fl = LU(fs.x < 0 ? 0 : fs.x);
      d.fx = Y(fs.y < 0 ? 0 : fs.y);
    };
    int2 fs_pre;  // slot of the set refilled at the next step
    auto bstep = [&](int k, USet& d) __attribute__((always_inline)) {
      const int m = n - 1 - k;
      const int4 mk = uni4(ba1), bk = uni4(bb1);
      ba1 = ba2;
      ba2 = ba3;
      ba3 = vload4(t.col_a, cl(k - 3, n - 1));
      bb1 = bb2;
      bb2 = vload4(t.col_b, cl(k - 2, n - 1));
      const int c0 = mk.x, ul = mk.z, bq0 = bk.x, bq1 = bk.y, bfar = bk.z;
      const T xk = ring[slot(k)] / d.pv;
      if (!finite_value(xk)) ok = false;
      if (g == 0) ring[slot(k)] = xk;
      Y(k) = xk;
      const int r = k - W - 1;
      if (r >= 0) {
        T enter = d.yr;
        if (bfar) {
          T acc = zero;
          if (d.fs.x >= 0) {
            const int j = d.fs.y;
            T xj = d.fx;  // x_j for j >= k + B was stored before this set was loaded
            if (j < k + B) xj = ring[slot(j)];
            acc = d.fl * xj;
          }
          strided(bq0, g, bq1, [&](int q) {  // entries beyond one per lane group
            const int j = t.bwd_col[q];
            T xj;
            if (j < k + B) xj = ring[slot(j)];
            else xj = Y(j);
            acc = acc + LU(t.bwd_pos[q]) * xj;
          });
          enter = d.yr - group_sum<TW>(acc);
        }
        if (g == m % G) ring[slot(r)] = enter;
      }
      {
        T rv[MS];
#pragma unroll
        for (int i = 0; i < MS; ++i) rv[i] = ring[slot(d.ur[i])];
#pragma unroll
        for (int i = 0; i < MS; ++i) {
          const int u = g + G * i;
          if (u < ul && k - d.ur[i] <= W) ring[slot(d.ur[i])] = rv[i] - d.uv[i] * xk;
        }
      }
      for (int b = G * MS; b < ul; b += G * MT) {  // long column: the rest, in batches
        int rj[MT];
        T uvt[MT], rv[MT];
#pragma unroll
        for (int u = 0; u < MT; ++u) {
          const int o = b + G * u + g;
          const int p = c0 + (o < ul ? o : 0);
          rj[u] = rowp[p];
          uvt[u] = LU(p);
        }
#pragma unroll
        for (int u = 0; u < MT; ++u) rv[u] = ring[slot(rj[u])];
#pragma unroll
        for (int u = 0; u < MT; ++u) {
          const int o = b + G * u + g;
          if (o < ul && k - rj[u] <= W) ring[slot(rj[u])] = rv[u] - uvt[u] * xk;
        }
      }
      // refill the set (after its last use, so the loads land in the same registers)
      const int2 fs = fs_pre;
      fs_pre = bslot_of(k - 3);
      u_issue(cl(k - 2, n - 1), uni4(ba2), d, fs);
      if ((m + 1) % B == 0) workgroup_fence();
    };
    u_issue(n - 1, uni4(ba1), A, bslot_of(n - 1));
    u_issue(cl(n - 2, n - 1), uni4(ba2), Bs, bslot_of(n - 2));
    fs_pre = bslot_of(n - 3);
    drain();  // (as before the forward loop)
    for (int k = n - 1; k >= 0; k -= 2) {
      bstep(k, A);
      if (k == 0) break;
      bstep(k - 1, Bs);
    }
  }
  // a zero pivot seen by any group of the scenario, or a non-finite solution, makes it Singular
  ok = all_groups<TW>(ok && !overflow);
  if (g == 0 && !ok && w.status[sc] == kRunning && (kFactor || sc_refines(w, sc, iter))) {
    w.status[sc] = kSingular;
    w.iters[sc] = iter + 1;
  }
}

// Wide tile: one wavefront per tile of 64 scenarios, lane = scenario.  Every index is the same
// for all lanes, every access is one contiguous 64-value row, and a lane only ever reads values of
// its own scenario, so no fences are needed.
//   LDS: xc[cap+1][64] (the column being factored) and ring[R][64] with R = W + 1: at column k the
//   forward ring holds rows [k, k + W]; row k + W + 1 enters in the slot of row k once y_k has
//   been read, as f minus its far L entries taken in ascending column order (the order of the
//   column-oriented substitution).  Back substitution mirrors this with U.
// Refactor of column k: stage A(:,k), run its update rounds (kRoundItems independent items
// x_t -= L(i,j) * x_u each), divide the L part by the pivot, write back.
// Memory: loads are issued ahead of use so that each waits for data that has been in flight for a
// column or a round: column metadata and far-entry slots two columns ahead; staged values, row
// indices, the entering row's f and its first kWideFar far values one column ahead; round items
// two rounds ahead and their L values one round ahead (rounds form one stream over all columns;
// the host keeps L(:,k-1) out of column k's first round, so those loads never read a column
// before it is written).
template <typename T, bool kFactor>
__device__ __forceinline__ void lu64_tile(const DevTopo& t, const DevWork& w, int tile, int cap, int iter) {
  constexpr int TW = kWideTile;
  constexpr int IPR = kRoundItems;
  // complex values (the linear initial guess, run once per batch) take the shallow pipeline
  constexpr bool kDeep = sizeof(T) <= 8;
  constexpr int MS = kDeep ? kWideStage : 4;
  constexpr int MF = kWideFar;
  const int s = threadIdx.x;
  if (!tile_running<TW>(w.status, tile)) return;
  if (!kFactor && !w.swept[tile]) return;  // (the mixed correction solve)
  const int sc = tile * TW + s;
  T* __restrict__ lu = reinterpret_cast<T*>(w.lu) + lu_at(0, t.nnz_lu, tile, TW) + s;
  const size_t pst = (size_t)lu_block(TW) * TW;  // elements between consecutive positions
  T* __restrict__ y = reinterpret_cast<T*>(w.y) + at<TW>(tile, t.n, 0) + s;
  const int* __restrict__ lrow = t.lu_row;
  const int2* __restrict__ uitems = t.uitems;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* xc = reinterpret_cast<T*>(smem_raw);
  T* ring = xc + (kFactor ? (size_t)(cap + 1) * TW : 0);
  const int n = t.n, W = t.window, R = W + 1, nl = t.nnz_lu;
  auto rs = [&](int r) { return (r & (R - 1)) * TW + s; };
  auto xs = [&](int o) { return o * TW + s; };
  auto cl = [](int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); };
  bool ok = true;
  const T zero = zero_value<T>();

  // ------------------------------------------------ forward pass (+ refactor)
  for (int r = 0; r <= W && r < n; ++r) ring[rs(r)] = y[(size_t)r * TW];
  // level 1 (two columns ahead): metadata {c0, len, ul, first round}, overflow range, far slots
  int4 m1 = t.col_a[0], m2 = t.col_a[cl(1, n - 1)];
  int4 q1 = t.col_f[0], q2 = t.col_f[cl(1, n - 1)];
  int2 fs1[MF], fs2[MF];
#pragma unroll
  for (int i = 0; i < MF; ++i) {
    fs1[i] = t.fslot[i];
    fs2[i] = t.fslot[(size_t)cl(1, n - 1) * MF + i];
  }
  // level 2 (one column ahead): staged values (factor: A(:,k) from offset 0; solve: L(:,k)),
  // L-part row indices, the entering row's f, far values
  T nsv[MS], nfl[MF], nfy[MF], nfr;
  int nlr[MS];
  // (column kk's metadata in m1, its far slots in fs1)
  auto lvl2_issue = [&](int kk) {
    const int c0 = m1.x, len = m1.y, ul = m1.z;
    const int f0 = kFactor ? 0 : ul + 1;
#pragma unroll
    for (int i = 0; i < MS; ++i) {
      const int o = f0 + i;
      nsv[i] = lu[(size_t)(c0 + (o < len ? o : 0)) * pst];
    }
#pragma unroll
    for (int i = 0; i < MS; ++i) {
      const int o = ul + 1 + i;
      nlr[i] = lrow[c0 + (o < len ? o : 0)];
    }
#pragma unroll
    for (int i = 0; i < MF; ++i) {
      nfl[i] = lu[(size_t)cl(fs1[i].x, nl - 1) * pst];
      nfy[i] = y[(size_t)cl(fs1[i].y, n - 1) * TW];
    }
    nfr = y[(size_t)cl(kk + W + 1, n - 1) * TW];
  };
  lvl2_issue(0);
  // round stream: items two rounds ahead, L values one round ahead
  const int n_rounds = kFactor ? t.uround[n] : 0;
  int2 i1[kDeep ? IPR : 1], i2[kDeep ? IPR : 1];
  T v1[kDeep ? IPR : 1];
  if (kFactor && kDeep) {
#pragma unroll
    for (int i = 0; i < IPR; ++i) {
      i1[i] = uitems[i];
      i2[i] = uitems[(size_t)cl(1, n_rounds - 1) * IPR + i];
    }
#pragma unroll
    for (int i = 0; i < IPR; ++i) v1[i] = lu[(size_t)(i1[i].y < 0 ? 0 : i1[i].y) * pst];
  }
  for (int k = 0; k < n; ++k) {
    // this column's data; the next columns' loads
    const int4 mk = m1, qk = q1;
    int2 cfs[MF];
    T cvs[MS], cfl[MF], cfy[MF];
    int clr[MS];
#pragma unroll
    for (int i = 0; i < MS; ++i) {
      cvs[i] = nsv[i];
      clr[i] = nlr[i];
    }
#pragma unroll
    for (int i = 0; i < MF; ++i) {
      cfs[i] = fs1[i];
      cfl[i] = nfl[i];
      cfy[i] = nfy[i];
      fs1[i] = fs2[i];
    }
    const T cfr = nfr;
    m1 = m2;
    q1 = q2;
    {
      const int k2 = cl(k + 2, n - 1);
      m2 = t.col_a[k2];
      q2 = t.col_f[k2];
#pragma unroll
      for (int i = 0; i < MF; ++i) fs2[i] = t.fslot[(size_t)k2 * MF + i];
    }
    if (k + 1 < n) lvl2_issue(k + 1);
    const int c0 = mk.x, len = mk.y, ul = mk.z;
    if (kFactor) {
      if (len > cap) {  // the host sizes cap to the longest column; kept as a guard
        ok = false;
        break;
      }
#pragma unroll
      for (int i = 0; i < MS; ++i)
        if (i < len) xc[xs(i)] = cvs[i];
      for (int o = MS; o < len; ++o) xc[xs(o)] = lu[(size_t)(c0 + o) * pst];
      const int r0 = mk.w, r1 = k + 1 < n ? m1.w : n_rounds;
      for (int rho = r0; rho < r1; ++rho) {
        int2 it[IPR];
        T lv[IPR];
        if constexpr (kDeep) {
#pragma unroll
          for (int i = 0; i < IPR; ++i) {
            it[i] = i1[i];
            lv[i] = i1[i].y < 0 ? zero : v1[i];
            i1[i] = i2[i];
          }
#pragma unroll
          for (int i = 0; i < IPR; ++i) v1[i] = lu[(size_t)(i1[i].y < 0 ? 0 : i1[i].y) * pst];
          const int rn = cl(rho + 2, n_rounds - 1);
#pragma unroll
          for (int i = 0; i < IPR; ++i) i2[i] = uitems[(size_t)rn * IPR + i];
        } else {
#pragma unroll
          for (int i = 0; i < IPR; ++i) {
            it[i] = uitems[(size_t)rho * IPR + i];
            lv[i] = it[i].y < 0 ? zero : lu[(size_t)(it[i].y < 0 ? 0 : it[i].y) * pst];
          }
        }
        if (it[0].y >= 0) {  // empty rounds (columns without updates) do nothing
          T xt[IPR], xu[IPR];
#pragma unroll
          for (int i = 0; i < IPR; ++i) {
            xt[i] = xc[xs(it[i].x & 0xffff)];
            xu[i] = xc[xs(it[i].x >> 16)];
          }
#pragma unroll
          for (int i = 0; i < IPR; ++i) xc[xs(it[i].x & 0xffff)] = xt[i] - lv[i] * xu[i];
        }
      }
      const T ukk = xc[xs(ul)];
      if (is_zero(ukk)) ok = false;
      for (int o = ul + 1; o < len; ++o) xc[xs(o)] = xc[xs(o)] / ukk;
      for (int o = 0; o < len; ++o) lu[(size_t)(c0 + o) * pst] = xc[xs(o)];
    }
    // forward substitution: y_k is final
    const T yk = ring[rs(k)];
    y[(size_t)k * TW] = yk;
    // the row entering the window after column k: f minus its far L entries, ascending columns
    const int rin = k + W + 1;
    if (rin < n) {
      T v = cfr;
#pragma unroll
      for (int i = 0; i < MF; ++i) {
        if (cfs[i].x >= 0) {
          const int p = cfs[i].x, j = cfs[i].y;
          T lj = cfl[i], yj = cfy[i];
          if (kFactor && j == k) lj = xc[xs(p - c0)];
          if (kFactor && j == k - 1) lj = lu[(size_t)p * pst];
          if (j == k) yj = yk;
          if (j == k - 1) yj = y[(size_t)j * TW];
          v = v - lj * yj;
        }
      }
      for (int q = qk.y; q < qk.z; ++q) {
        const int p = t.fwd_pos[q], j = t.fwd_col[q];
        T lj;
        if (kFactor && j == k) lj = xc[xs(p - c0)];
        else lj = lu[(size_t)p * pst];
        const T yj = j == k ? yk : y[(size_t)j * TW];
        v = v - lj * yj;
      }
      ring[rs(rin)] = v;
    }
    // push L(:,k) into the window rows (farther rows are gathered when they enter)
#pragma unroll
    for (int i = 0; i < MS; ++i) {
      const int o = ul + 1 + i;
      if (o < len) {
        const int r = clr[i];
        const T lv = kFactor ? xc[xs(o)] : cvs[i];
        if (r - k <= W) ring[rs(r)] = ring[rs(r)] - lv * yk;
      }
    }
    for (int o = ul + 1 + MS; o < len; ++o) {
      const int r = lrow[c0 + o];
      const T lv = kFactor ? xc[xs(o)] : lu[(size_t)(c0 + o) * pst];
      if (r - k <= W) ring[rs(r)] = ring[rs(r)] - lv * yk;
    }
  }

  // ------------------------------------------------ back substitution
  if (ok) {
    for (int r = n - 1; r >= 0 && r >= n - 1 - W; --r) ring[rs(r)] = y[(size_t)r * TW];
    // level 1: metadata, overflow range and far slots of column k - 2; level 2: U values and rows,
    // pivot, the entering row's y and far values of column k - 1
    int4 b1 = t.col_a[n - 1], b2 = t.col_a[cl(n - 2, n - 1)];
    int4 e1 = t.col_b[n - 1], e2 = t.col_b[cl(n - 2, n - 1)];
    int2 bs1[MF], bs2[MF];
#pragma unroll
    for (int i = 0; i < MF; ++i) {
      bs1[i] = t.bslot[(size_t)(n - 1) * MF + i];
      bs2[i] = t.bslot[(size_t)cl(n - 2, n - 1) * MF + i];
    }
    T nuv[MS], npv, nbl[MF], nbx[MF], nyr;
    int nur[MS];
    // (column kk's metadata in b1, its far slots in bs1)
    auto b2_issue = [&](int kk) {
      const int c0 = b1.x, ul = b1.z;
#pragma unroll
      for (int i = 0; i < MS; ++i) {
        const int p = c0 + (i < ul ? i : 0);
        nuv[i] = lu[(size_t)p * pst];
        nur[i] = lrow[p];
      }
      npv = lu[(size_t)(c0 + ul) * pst];
#pragma unroll
      for (int i = 0; i < MF; ++i) {
        nbl[i] = lu[(size_t)cl(bs1[i].x, nl - 1) * pst];
        nbx[i] = y[(size_t)cl(bs1[i].y, n - 1) * TW];
      }
      nyr = y[(size_t)cl(kk - W - 1, n - 1) * TW];
    };
    b2_issue(n - 1);
    for (int k = n - 1; k >= 0; --k) {
      const int4 mk = b1, ek = e1;
      int2 cbs[MF];
      T cuv[MS], cbl[MF], cbx[MF];
      int cur_[MS];
#pragma unroll
      for (int i = 0; i < MS; ++i) {
        cuv[i] = nuv[i];
        cur_[i] = nur[i];
      }
#pragma unroll
      for (int i = 0; i < MF; ++i) {
        cbs[i] = bs1[i];
        cbl[i] = nbl[i];
        cbx[i] = nbx[i];
        bs1[i] = bs2[i];
      }
      const T cpv = npv, cyr = nyr;
      b1 = b2;
      e1 = e2;
      {
        const int k2 = cl(k - 2, n - 1);
        b2 = t.col_a[k2];
        e2 = t.col_b[k2];
#pragma unroll
        for (int i = 0; i < MF; ++i) bs2[i] = t.bslot[(size_t)k2 * MF + i];
      }
      if (k > 0) b2_issue(k - 1);
      const int c0 = mk.x, ul = mk.z;
      const T xk = ring[rs(k)] / cpv;
      if (!finite_value(xk)) ok = false;
      y[(size_t)k * TW] = xk;
      const int rin = k - W - 1;
      if (rin >= 0) {
        T v = cyr;
#pragma unroll
        for (int i = 0; i < MF; ++i) {
          if (cbs[i].x >= 0) {
            const int j = cbs[i].y;
            T xj = cbx[i];
            if (j == k) xj = xk;
            if (j == k + 1) xj = y[(size_t)j * TW];
            v = v - cbl[i] * xj;
          }
        }
        for (int q = ek.x; q < ek.y; ++q) {
          const int j = t.bwd_col[q];
          const T xj = j == k ? xk : y[(size_t)j * TW];
          v = v - lu[(size_t)t.bwd_pos[q] * pst] * xj;
        }
        ring[rs(rin)] = v;
      }
#pragma unroll
      for (int i = 0; i < MS; ++i) {
        if (i < ul) {
          const int r = cur_[i];
          if (k - r <= W) ring[rs(r)] = ring[rs(r)] - cuv[i] * xk;  // farther rows: gathered on entry
        }
      }
      for (int u = MS; u < ul; ++u) {
        const int r = lrow[c0 + u];
        if (k - r <= W) ring[rs(r)] = ring[rs(r)] - lu[(size_t)(c0 + u) * pst] * xk;
      }
    }
  }
  if (!ok && w.status[sc] == kRunning && (kFactor || sc_refines(w, sc, iter))) {
    w.status[sc] = kSingular;
    w.iters[sc] = iter + 1;
  }
}

// One launch per batch: one wavefront per tile (XCD-aware numbering).
// Refactor by parts, then the forward and back substitutions by levels; one wavefront per tile.
//   LDS: two part buffers xc[2][cap+2][S] (part p in buffer p & 1: the previous part stays
//   readable for the updates that use its L values; S = part_row_stride) while factoring, then
//   ring[R][TW+1] for the solves from offset 0.
// A part is a set of columns of one elimination level (independent of each other), laid one after
// the other in its buffer: stage A at their entries, run the part's update rounds (the chunk stream
// of the column-by-column kernel, host-scheduled), divide each L entry by its column's pivot, write
// back.  The first kPartRegs * G entries of a part are staged through registers loaded a part ahead
// (their positions two parts ahead); the rest are staged by the part's first rounds, whose values come
// with the round stream's loads.
// TR / TWR: the refactor's value type and lane-group width when they differ from the substitutions'
// (quad tiles: the refactor runs 32 lane groups of two lanes, four scenarios per lane, the
// substitutions 16 groups of four lanes, two scenarios per lane; both address the same tile layout).
template <typename T, int TW, bool kFactor, typename TR = T, int TWR = TW>
__device__ __forceinline__ void lu_tile_parts(const DevTopo& t, const DevWork& w, int tile, int cap, int iter) {
  constexpr int G = kLanes / TWR;         // the refactor's lane groups
  constexpr int VR = lane_pack<TR>::V;    // scenarios per refactor lane
  constexpr int S = part_row_stride(TWR, VR);
  constexpr int NB = part_chunk_rounds(G);
  constexpr int MP = kPartRegs;
  constexpr int MT = 4;  // long-part entries per lane group and batch
  const int W = t.window;
  const int R = W + kRingGap + 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // LDS (lu_parts_lds_bytes): while factoring, the two part buffers and the descriptors of a
  // part's entries past the registered ones ("tail"), two parts' worth; the solve ring from offset 0
  // once the refactor is done (the phases follow each other in the wavefront's program order)
  TR* xcb = reinterpret_cast<TR*>(smem_raw);
  const int xstride = (cap + 2) * S;  // rows: cap entries, the empty-slot row, the row holding -1
  T* ring = reinterpret_cast<T*>(smem_raw);
  int2* tdsc = reinterpret_cast<int2*>(smem_raw + (2 * (size_t)xstride * sizeof(TR) + 15) / 16 * 16);
  constexpr int V = lane_pack<T>::V;  // scenarios per substitution lane (2 in pair tiles)
  static_assert(TW * V == TWR * VR && TW * sizeof(T) == TWR * sizeof(TR), "both phases cover the tile's values");
  const int g = grp<TWR>(), s = sub<TWR>();  // the refactor's lane group and lane within it
  const int gs = grp<TW>(), ss = sub<TW>();  // the substitutions'
  if (!tile_running_v<TW, V>(w.status, tile)) return;
  if (!kFactor && !w.swept[tile]) return;  // (the mixed correction solve)
  const int nl = t.nnz_lu;
  const char* __restrict__ lbase = reinterpret_cast<const char*>(reinterpret_cast<T*>(w.lu) + lu_at(0, t.nnz_lu, tile, TW));
  T* __restrict__ yb = reinterpret_cast<T*>(w.y) + (size_t)tile * t.n * TW;
  const unsigned pstb = (unsigned)(lu_block(TW) * TW * sizeof(T));
  // a uniform base plus a 32-bit lane offset (the scalar-base addressing form of the loads)
  const unsigned loff = (unsigned)(s * sizeof(TR));
  auto LU = [&](int p) -> TR& { return *reinterpret_cast<TR*>(const_cast<char*>(lbase) + ((unsigned)p * pstb + loff)); };
  int bad = 0;   // the refactor lane's scenarios with a zero pivot (bit per scenario)
  int sbad = 0;  // the substitution lane's scenarios with a non-finite solution
  bool overflow = false;
  auto cl = [](int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); };
  if (kFactor) {
    const int np = t.n_parts;
    const int* __restrict__ ppos = t.pent_pos;
    const int* __restrict__ pdiv = t.pent_div;
    // staged entries o = g + G * i of a part: c* = the part being factored (values in registers),
    // x* = the next part (values in flight), y* = the one after (positions in flight)
    int cpos[MP], cdiv[MP], xpos[MP], xdiv[MP], ypos[MP], ydiv[MP];
    // tail descriptors (entries G MP + t, t = lane + 64 i < kPartTail) of the next part, in flight
    constexpr int MDT = kPartTail / kLanes;
    int tpos[MDT], tdiv[MDT];
    auto tail_issue = [&](const int4& m) {
#pragma unroll
      for (int i = 0; i < MDT; ++i) {
        const int o = G * MP + lane_id() + kLanes * i;
        const int e = m.x + (o < m.y ? o : 0);
        tpos[i] = ppos[e];
        tdiv[i] = pdiv[e];
      }
    };
    const int ts = part_tail(cap, TWR);  // tail descriptors per part held in LDS
    auto tail_store = [&](int buf, int len) {
#pragma unroll
      for (int i = 0; i < MDT; ++i) {
        const int tt = lane_id() + kLanes * i;
        if (tt < ts && G * MP + tt < len) tdsc[buf * ts + tt] = make_int2(tpos[i], tdiv[i]);
      }
    };
    // descriptor of entry o >= G MP of part p (pb: its first entry)
    // (lds_only: the caller knows o is within the LDS copy; else a global read past it.  One pointer
    // chosen between the two would make a flat access, whose wait also waits for every load in
    // flight.)
    auto tail_desc = [&](int p, int pb, int o, auto lds_only) -> int2 {
      const int tt = o - G * MP;
      if constexpr (decltype(lds_only)::value) return tdsc[(p & 1) * ts + tt];
      int2 d = tdsc[(p & 1) * ts + (tt < ts ? tt : 0)];
      if (tt >= ts) d = make_int2(global_int(ppos + pb + o), global_int(pdiv + pb + o));
      return d;
    };
    TR cval[MP], xval[MP];
    auto ent_issue = [&](const int4& m, int (&pp)[MP], int (&dd)[MP]) {
#pragma unroll
      for (int i = 0; i < MP; ++i) {
        const int o = g + G * i;
        const int e = m.x + (o < m.y ? o : 0);
        pp[i] = ppos[e];
        dd[i] = pdiv[e];
      }
    };
    auto val_issue_p = [&](const int (&pp)[MP], TR (&vv)[MP]) {
#pragma unroll
      for (int i = 0; i < MP; ++i) vv[i] = LU(pp[i] & ~kStageZero);
    };
    // The update rounds of all parts form one stream of chunks of NB rounds (part by part, every
    // part at least kPartMinChunks chunks).  The item positions of chunk gc are loaded kPartAhead
    // chunks before it runs, its values and words kPartMinChunks chunks before; three register sets
    // take turns by the chunk's index mod 3, so the loop copies no register a load is filling.
    static_assert(NB % kJointRounds == 0, "a chunk holds whole joint rounds");
    static_assert(kPartMinChunks == 2 && kPartAhead == 4, "the register roles below assume these");
    int xs[3][NB], ws[3][NB];
    TR vs[3][NB];
    const int* __restrict__ xsb = t.part_x + g * NB;
    const int* __restrict__ wsb = t.part_w + g * NB;
    auto ints_issue = [&](const int* __restrict__ base, int chunk, int (&d)[NB]) {
      const int4* q = reinterpret_cast<const int4*>(base + (size_t)(unsigned)chunk * (G * NB));
#pragma unroll
      for (int i = 0; i < NB / 4; ++i) {
        const int4 v = q[i];
        d[4 * i] = v.x;
        d[4 * i + 1] = v.y;
        d[4 * i + 2] = v.z;
        d[4 * i + 3] = v.w;
      }
    };
    auto vals_issue = [&](const int (&x)[NB], TR (&d)[NB]) {
#pragma unroll
      for (int r = 0; r < NB; ++r) d[r] = LU(x[r]);
    };
    // chunk gc in register role E: the loads of the chunks ahead, then its rounds, kJointRounds at a
    // time: no item of a joint round reads or writes another's target, so all their LDS reads go
    // before their writes.  Item word: target | source << 10 | (LDS offset of a previous-part L value)
    // << 20 | previous-part << 31.  A chunk with rleft > 0 runs all of its NB rounds: the rounds past
    // the part's last one hold empty items (the unused row past the part updated by itself), so the
    // rounds need no branch between them and each round's address arithmetic is issued during the
    // previous round's LDS wait (2,059 parts on the 10,000-bus grid: 197,984 rounds run for 182,052).
    auto step = [&](auto role, int gc, int rleft, TR* xc, const TR* xprev) __attribute__((always_inline)) {
      constexpr int E = decltype(role)::value, E1 = (E + 1) % 3, E2 = (E + 2) % 3;
      vals_issue(xs[E2], vs[E2]);                    // values of chunk gc + 2
      ints_issue(xsb, gc + kPartAhead, xs[E1]);      // positions of chunk gc + 4
      ints_issue(wsb, gc + kPartMinChunks, ws[E2]);  // words of chunk gc + 2
      if (rleft <= 0) return;
      // The L values of the previous part (final in its LDS buffer before this part's rounds start)
      // replace the loaded ones first, in the chunks that have any (2% of the items, 21% of the
      // chunks on the 10,000-bus grid), so no round holds a conditional LDS read.
      int flagged = 0;
#pragma unroll
      for (int r = 0; r < NB; ++r) flagged |= ws[E][r];
      if (__ballot(flagged < 0) != 0) {
#pragma unroll
        for (int r = 0; r < NB; ++r) {
          const int wd = ws[E][r];
          const TR pv = xprev[((wd >> 20) & 0x3ff) * S + s];
          if (wd < 0) vs[E][r] = pv;
        }
      }
      // LDS byte addresses as one multiply-add each: the row size as an opaque scalar keeps the
      // compiler from turning a power-of-two multiply into a shift and a mask
      unsigned rowb = S * sizeof(TR);
      asm volatile("" : "+s"(rowb));
      const unsigned lo = (unsigned)(reinterpret_cast<const char*>(xc) - reinterpret_cast<const char*>(smem_raw)) +
                          (unsigned)(s * sizeof(TR));
      auto X = [&](unsigned row) -> TR& { return *reinterpret_cast<TR*>(smem_raw + (__umul24(row, rowb) + lo)); };
#pragma unroll
      for (int r = 0; r < NB; r += kJointRounds) {
        unsigned to[kJointRounds];
        TR nv[kJointRounds];
#pragma unroll
        for (int h = 0; h < kJointRounds; ++h) {
          const unsigned wd = (unsigned)ws[E][r + h];
          to[h] = wd & 0x3ffu;
          nv[h] = fnma(vs[E][r + h], X(__builtin_amdgcn_ubfe(wd, 10, 10)), X(to[h]));
        }
#pragma unroll
        for (int h = 0; h < kJointRounds; ++h) X(to[h]) = nv[h];
      }
    };
    int4 m1 = vload4(t.part_a, cl(1, np - 1)), m2 = vload4(t.part_a, cl(2, np - 1));
    {
      const int4 m0 = uni4(vload4(t.part_a, 0));
      ent_issue(m0, cpos, cdiv);
      ent_issue(uni4(m1), xpos, xdiv);
      val_issue_p(cpos, cval);
    }
    // the stream's first chunks: positions 0-3, words 0-1, values 0-1
    ints_issue(xsb, 0, xs[0]);
    ints_issue(xsb, 1, xs[1]);
    ints_issue(xsb, 2, xs[2]);
    ints_issue(wsb, 0, ws[0]);
    ints_issue(wsb, 1, ws[1]);
    drain();
    vals_issue(xs[0], vs[0]);
    vals_issue(xs[1], vs[1]);
    ints_issue(xsb, 3, xs[0]);
    drain();
    int4 mk = uni4(vload4(t.part_a, 0));
    tail_issue(mk);
    // the row after the empty-slot row holds -1 in both part buffers (staging items multiply by it)
    xcb[(cap + 1) * S + s] = minus_one<TR>();
    xcb[xstride + (cap + 1) * S + s] = minus_one<TR>();
    int role = 0;  // the running chunk's index mod 3
    for (int p = 0; p < np; ++p) {
      const int pb = mk.x, len = mk.y, ch0 = mk.z, nr = mk.w;
      const int4 mx = uni4(m1);
      m1 = m2;
      m2 = vload4(t.part_a, cl(p + 3, np - 1));
      TR* xc = xcb + (p & 1) * xstride;
      const TR* xprev = xcb + ((p & 1) ^ 1) * xstride;
      if (len > cap) {  // the host sizes cap to the largest part; kept as a guard
        overflow = true;
        break;
      }
      // this part's tail descriptors (loaded a part ago) into LDS; the next part's in flight
      tail_store(p & 1, len);
      memory_order_point();  // the descriptors in LDS before other lanes read them
      tail_issue(mx);
      // stage the registered entries; the others start at zero (their staging items add the values
      // of those that are not structural fill)
#pragma unroll
      for (int i = 0; i < MP; ++i) {
        const int o = g + G * i;
        if (o < len) xc[o * S + s] = (cpos[i] & kStageZero) ? zero_value<TR>() : cval[i];
      }
      for (int o = G * MP + g; o < len; o += G) xc[o * S + s] = zero_value<TR>();
      // loads for the next parts
      val_issue_p(xpos, xval);
      {
        const int4 my = uni4(m1);
        ent_issue(my, ypos, ydiv);
      }
      memory_order_point();  // the staged values before the rounds read them
      const int nch = (nr + NB - 1) / NB > kPartMinChunks ? (nr + NB - 1) / NB : kPartMinChunks;
      for (int c = 0;;) {
        if (role == 0) {
          step(std::integral_constant<int, 0>{}, ch0 + c, nr - c * NB, xc, xprev);
          role = 1;
          if (++c == nch) break;
        }
        if (role == 1) {
          step(std::integral_constant<int, 1>{}, ch0 + c, nr - c * NB, xc, xprev);
          role = 2;
          if (++c == nch) break;
        }
        if (role == 2) {
          step(std::integral_constant<int, 2>{}, ch0 + c, nr - c * NB, xc, xprev);
          role = 0;
          if (++c == nch) break;
        }
      }
      memory_order_point();  // the rounds' writes before the division reads
      // divide the L entries by their columns' pivots, write back
#pragma unroll
      for (int i = 0; i < MP; ++i) {
        const int o = g + G * i;
        if (o < len) {
          TR v = xc[o * S + s];
          const int d = cdiv[i];
          if (d >= 0) v = v / xc[d * S + s];
          if (d == -2) bad |= zero_mask(v);
          cval[i] = v;
        }
      }
      memory_order_point();  // every pivot read before the divided values are stored
#pragma unroll
      for (int i = 0; i < MP; ++i) {
        const int o = g + G * i;
        if (o < len) {
          if (cdiv[i] >= 0) xc[o * S + s] = cval[i];
          LU(cpos[i] & ~kStageZero) = cval[i];
        }
      }
      // entries past the registered ones, MT per lane group and batch: a fixed number of batches
      // (their descriptors sit in LDS), so the loads and stores of a part transition are the same
      // on every path and the waits of the round stream stay counted; longer parts finish in a
      // loop of their own, drained before and after
      auto tail_batch = [&](int b, auto lds_only) __attribute__((always_inline)) {
        int pp[MT], dd[MT];
        TR tv[MT];
#pragma unroll
        for (int u = 0; u < MT; ++u) {
          const int o = b + G * u + g;
          const int2 d = tail_desc(p, pb, o < len ? o : G * MP, lds_only);
          pp[u] = d.x & ~kStageZero;
          dd[u] = d.y;
        }
#pragma unroll
        for (int u = 0; u < MT; ++u) {
          const int o = b + G * u + g;
          tv[u] = xc[(o < len ? o : 0) * S + s];
          if (dd[u] >= 0) tv[u] = tv[u] / xc[dd[u] * S + s];
          if (o < len && dd[u] == -2) bad |= zero_mask(tv[u]);
        }
        memory_order_point();
#pragma unroll
        for (int u = 0; u < MT; ++u) {
          const int o = b + G * u + g;
          if (o < len) {
            if (dd[u] >= 0) xc[o * S + s] = tv[u];
            LU(pp[u]) = tv[u];
          }
        }
      };
      // (the batches below cover at most kPartTail entries past the registered ones, all of them
      // within the LDS copy of min(cap - G MP, kPartTail) descriptors)
      constexpr int kTailBatches = kPartTail / (G * MT);
      for (int b = G * MP; b < len && b < G * MP + kTailBatches * G * MT; b += G * MT) tail_batch(b, std::true_type{});
      if (len > G * MP + kTailBatches * G * MT) {
        drain();
        for (int b = G * MP + kTailBatches * G * MT; b < len; b += G * MT) tail_batch(b, std::false_type{});
        drain();
      }
      // the registered sets rotate
#pragma unroll
      for (int i = 0; i < MP; ++i) {
        cpos[i] = xpos[i];
        cdiv[i] = xdiv[i];
        cval[i] = xval[i];
        xpos[i] = ypos[i];
        xdiv[i] = ydiv[i];
      }
      mk = mx;
      memory_order_point();  // this part's LDS and memory writes before the next part's reads
    }
  }
  if (!overflow) {
    memory_order_point();
    level_solve<T, TW, false>(t.fwd_items, gs, t.fwd_steps, R, nl, lbase, pstb, reinterpret_cast<char*>(yb),
                             (unsigned)(ss * sizeof(T)), ring + ss, sbad);
    memory_order_point();
    level_solve<T, TW, true>(t.back_items, gs, t.back_steps, R, nl, lbase, pstb, reinterpret_cast<char*>(yb),
                             (unsigned)(ss * sizeof(T)), ring + ss, sbad);
  }
  // a zero pivot seen by any group of the scenario, or a non-finite solution, makes it Singular
  if constexpr (TW == TWR) {
    bad = any_groups<TW>(overflow ? (1 << V) - 1 : (bad | sbad));
    if (g == 0) {
#pragma unroll
      for (int e = 0; e < V; ++e) {
        const int sc = (tile * TW + s) * V + e;
        if ((bad >> e & 1) && w.status[sc] == kRunning && (kFactor || sc_refines(w, sc, iter))) {
          w.status[sc] = kSingular;
          w.iters[sc] = iter + 1;
        }
      }
    }
  } else {
    // the two phases map lanes to scenarios differently: the flags in the tile's scenario order
    const int lm = overflow ? (1 << (TW * V)) - 1 : ((bad << (s * VR)) | (sbad << (ss * V)));
    int tm = 0;
#pragma unroll
    for (int e = 0; e < TW * V; ++e) tm |= __ballot((lm >> e) & 1) != 0ull ? 1 << e : 0;
    const int l = lane_id();
    if (l < TW * V) {
      const int sc = tile * TW * V + l;
      if ((tm >> l & 1) && w.status[sc] == kRunning && (kFactor || sc_refines(w, sc, iter))) {
        w.status[sc] = kSingular;
        w.iters[sc] = iter + 1;
      }
    }
  }
}

template <typename T, int TW, bool kFactor>
__global__ __launch_bounds__(64) void k_lu(DevTopo t, DevWork w, int cap, int iter) {
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  if constexpr (TW == kWideTile) lu64_tile<T, kFactor>(t, w, tile, cap, iter);
  else lu_tile<T, TW, kFactor>(t, w, tile, cap, iter);
}

template <typename T, int TW, bool kFactor, typename TR = T, int TWR = TW>
__global__ __launch_bounds__(64) void k_lu_parts(DevTopo t, DevWork w, int cap, int iter) {
  lu_tile_parts<T, TW, kFactor, TR, TWR>(t, w, xcd_tile(blockIdx.x, gridDim.x), cap, iter);
}

// One persistent launch over the tiles of several requests (topologies of any size): each
// wavefront takes (request, tile) items, heaviest first, from a device counter until the list is
// exhausted, so every wavefront reaches the end of the list.  Each request's iteration number is
// read from its own counter.
template <typename T, int TW, bool kFactor>
__global__ __launch_bounds__(64) void k_lu_queue(const DevTopo* __restrict__ topos, const DevWork* __restrict__ works,
                                                 const int2* __restrict__ items, int n_items, int* __restrict__ counter) {
  for (;;) {
    int it = 0;
    if (threadIdx.x == 0) it = atomicAdd(counter, 1);
    // lane 0 is the first active lane: the item index becomes a scalar, and with it every
    // topology field (scalar loads), as in the one-batch kernel
    it = __builtin_amdgcn_readfirstlane(it);
    if (it >= n_items) break;
    const int2 item = items[it];
    const DevTopo& t = topos[item.x];
    const DevWork& w = works[item.x];
    const int cap = t.max_col < 1 ? 1 : t.max_col;
    const int iter = w.iter_dev ? *w.iter_dev : 0;
    if constexpr (TW == kWideTile) {
      lu64_tile<T, kFactor>(t, w, item.y, cap, iter);
    } else {
      bool done = false;
      if constexpr (TW == 8 && (std::is_same<T, float>::value || std::is_same<T, double>::value)) {
        // 8-scenario tiles staged as pair tiles (4-lane groups, two scenarios per lane), FP32
        // refactors as quad tiles where the 32-group parts exist, as in run_lu
        using P = pair2<T>;
        if (t.part_a && t.groups == kLanes / 4) {
          if constexpr (std::is_same<T, float>::value) {
            if (t.part_groups == kLanes / 2) {
              lu_tile_parts<P, 4, kFactor, pair2<P>, 2>(t, w, item.y, cap, iter);
              done = true;
            }
          }
          if (!done && t.part_groups == kLanes / 4) {
            lu_tile_parts<P, 4, kFactor>(t, w, item.y, cap, iter);
            done = true;
          }
        }
      }
      if (!done) {
        if (t.part_a && t.part_groups == kLanes / TW) lu_tile_parts<T, TW, kFactor>(t, w, item.y, cap, iter);
        else lu_tile<T, TW, kFactor>(t, w, item.y, cap, iter);
      }
    }
    __syncthreads();  // the next item reuses the LDS
  }
}

__global__ void k_zero_ints(int* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// End of a Newton step of a request whose iteration number lives on the device.
__global__ void k_iter_advance(int* __restrict__ iter_dev) {
  if (threadIdx.x == 0) *iter_dev += 1;
}

// Mixed precision, refinement step 1: dx1 (FP64 copy of the FP32 solution) and
// dV_i = V_i * (d|V|_i / |V|_i + j dtheta_i) per bus.
template <typename T, int TW>
__global__ __launch_bounds__(256) void k_dv(DevTopo t, DevWork w, double* __restrict__ dvr,
                                            double* __restrict__ dvi, int iter) {
  const int tile = blockIdx.y, s = sub<TW>();
  if (!tile_running<TW>(w.status, tile)) return;
  // the tile's sweep decision, fixed here for the rest of the step
  const bool rf = tile_refines<TW>(w, tile, iter);
  if (blockIdx.x == 0 && threadIdx.x == 0) w.swept[tile] = rf ? 1 : 0;
  if (!rf) return;
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

// Mixed precision, refinement step 2: r = f - J*dx in FP64, stored (in LU precision) as the
// right-hand side of the correction solve.
template <typename T, int TW>
__global__ __launch_bounds__(256) void k_resid(DevTopo t, DevWork w, const double* __restrict__ dvr,
                                               const double* __restrict__ dvi, int iter) {
  const int tile = blockIdx.y, s = sub<TW>();
  if (!tile_running<TW>(w.status, tile) || !w.swept[tile]) return;
  const size_t bb = at<TW>(tile, t.n_bus, 0) + s;
  const size_t ub = at<TW>(tile, t.n, 0) + s;
  T* __restrict__ y = reinterpret_cast<T*>(w.y);
  const Outage og = load_outage(t, w, tile * TW + s);
  for (int i = elem_first<TW>(); i < t.n_bus; i += elem_stride<TW>()) {
    const int rp = t.rpos_p[i], rq = t.rpos_q[i];
    if (rp < 0 && rq < 0) continue;
    double ar = 0.0, ai = 0.0;
    const int p2 = t.y_rowptr[i + 1];
    for (int p = t.y_rowptr[i]; p < p2; ++p) {
      const size_t oj = bb + (size_t)t.y_col[p] * TW;
      double g, b;
      y_at(t, og, p, g, b);
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
__global__ __launch_bounds__(256) void k_update(DevTopo t, DevWork w, int iter) {
  const int tile = blockIdx.y, s = sub<TW>();
  if (!tile_running<TW>(w.status, tile)) return;
  // mixed: when the tile ran the sweep, y holds the correction and dx the first solution (exactly
  // the FP32 step in FP64); a scenario that does not refine takes dx alone, else y (the step).
  // The sweep decision is the one k_dv recorded: a scenario the correction solve made Singular
  // does not change what its tile-mates take.
  const bool swept = kMixed && w.swept[tile] != 0;
  const int sc = tile * TW + s;
  const bool own = swept && sc_refines(w, sc, iter);
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
      double d = swept ? w.dx[ub + (size_t)xa * TW] : (double)y[ub + (size_t)xa * TW];
      if (own) d += (double)y[ub + (size_t)xa * TW];
      a += d;
    }
    if (xm >= 0) {
      double d = swept ? w.dx[ub + (size_t)xm * TW] : (double)y[ub + (size_t)xm * TW];
      if (own) d += (double)y[ub + (size_t)xm * TW];
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
  const Outage og = load_outage(t, w, sc);
  for (int i = elem_first<TW>(); i < t.n_bus; i += elem_stride<TW>()) {
    const int ty = t.bus_type[i];
    if (ty != 1 && ty != 2) continue;
    double ir = 0.0, ii = 0.0;
    const int p2 = t.y_rowptr[i + 1];
    for (int p = t.y_rowptr[i]; p < p2; ++p) {
      const size_t oj = bb + (size_t)t.y_col[p] * TW;
      double g, b;
      y_at(t, og, p, g, b);
      const double vr = w.vr[oj], vi = w.vi[oj];
      ir += g * vr - b * vi;
      ii += g * vi + b * vr;
    }
    const size_t o = bb + (size_t)i * TW;
    const double vr = w.vr[o], vi = w.vi[o];
    double pe = w.ps[o], qe = w.qs[o];
    effective_injection(t, i, w.vm[o], pe, qe);
    double mx = fabs(pe - (vr * ir + vi * ii));
    if (ty == 2) mx = fmax(mx, fabs(qe - (vi * ir - vr * ii)));
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
  zd* out = lu + lu_at(0, lg.nnz_lu, tile, TW) + s;
  const size_t pst = (size_t)lu_block(TW) * TW;
  const Outage og = load_outage(t, w, tile * TW + s);
  for (int p = elem_first<TW>(); p < lg.nnz_lu; p += elem_stride<TW>()) {
    zd v{0.0, 0.0};
    const int e = lg.src[p];
    if (e >= 0) y_at(t, og, e, v.re, v.im);
    const int d = lg.dbus[p];
    if (d >= 0) {
      const size_t o = bb + (size_t)d * TW;
      double pe = w.ps[o], qe = w.qs[o];
      effective_injection(t, d, w.vm[o], pe, qe);
      v.re += -pe;  // -conj(P + jQ) = -P + jQ
      v.im += qe;
    }
    out[(size_t)p * pst] = v;
  }
}

template <int TW>
__global__ __launch_bounds__(256) void k_lg_rhs(DevTopo t, DevLg lg, DevWork w, zd* __restrict__ y) {
  const int tile = blockIdx.y, s = sub<TW>();
  const size_t bb = at<TW>(tile, t.n_bus, 0) + s;
  zd* out = y + at<TW>(tile, lg.m, 0) + s;
  const Outage og = load_outage(t, w, tile * TW + s);
  for (int r = elem_first<TW>(); r < lg.m; r += elem_stride<TW>()) {
    zd b{0.0, 0.0};
    for (int q = lg.rhs_ptr[r]; q < lg.rhs_ptr[r + 1]; ++q) {
      const int e = lg.rhs_e[q];
      const size_t oj = bb + (size_t)lg.rhs_j[q] * TW;
      double sn, cs;
      sincos(w.va[oj], &sn, &cs);
      const zd u{w.vm[oj] * cs, w.vm[oj] * sn};
      zd ye;
      y_at(t, og, e, ye.re, ye.im);
      b = b - ye * u;
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

// Resident warm start (src/solver.rs:395-409 keeps a topology's state between solves): every
// scenario starts from warm[0..n_bus) (|V|) and warm[n_bus..2 n_bus) (theta) when *ok is set.
template <int TW>
__global__ __launch_bounds__(256) void k_warm_start(DevTopo t, DevWork w, const double* __restrict__ warm,
                                                    const int* __restrict__ ok) {
  if (!*ok) return;
  const int tile = blockIdx.y, s = sub<TW>();
  const size_t bb = at<TW>(tile, t.n_bus, 0) + s;
  for (int i = elem_first<TW>(); i < t.n_bus; i += elem_stride<TW>()) {
    w.vm[bb + (size_t)i * TW] = warm[i];
    w.va[bb + (size_t)i * TW] = warm[t.n_bus + i];
  }
}

// The solution of the first converged scenario (if any) becomes the resident warm start.  One
// workgroup: the first converged scenario by a minimum over the workgroup, then its buses copied out
// of the tile layout.
template <int TW>
__global__ __launch_bounds__(256) void k_keep_warm(DevTopo t, DevWork w, double* __restrict__ warm,
                                                   int* __restrict__ ok) {
  __shared__ int first;
  if (threadIdx.x == 0) first = INT_MAX;
  __syncthreads();
  int mine = INT_MAX;
  for (int sc = threadIdx.x; sc < w.n_scen; sc += blockDim.x)
    if (w.status[sc] == kConverged) {
      mine = sc;
      break;
    }
  if (mine != INT_MAX) atomicMin(&first, mine);
  __syncthreads();
  const int sc = first;
  if (sc == INT_MAX) return;
  const size_t bb = at<TW>(sc / TW, t.n_bus, 0) + sc % TW;
  for (int i = threadIdx.x; i < t.n_bus; i += blockDim.x) {
    warm[i] = w.vm[bb + (size_t)i * TW];
    warm[t.n_bus + i] = w.va[bb + (size_t)i * TW];
  }
  if (threadIdx.x == 0) *ok = 1;
}

__global__ void k_finish(DevWork w) {
  const int sc = blockIdx.x * blockDim.x + threadIdx.x;
  if (sc >= w.n_tiles * w.tile) return;
  if (w.status[sc] == kRunning) {
    w.status[sc] = kMaxIter;
    w.iters[sc] = w.max_iter;
  }
}

// Copies a compact vector into / out of scenario 0 of an interleaved array (elements `stride` apart).
__global__ void k_lane0(const double* __restrict__ src, double* __restrict__ dst, int count, int to_tile, int stride) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= count) return;
  if (to_tile) dst[(size_t)e * stride] = src[e];
  else dst[e] = src[(size_t)e * stride];
}

// ---------------------------------------------------------------- allocation-free system solve
// Device memory of gx_sys_factor_and_solve_values_noalloc: the packed L+U values, the right-hand
// side and the status of the one-scenario tile.  Every GPU that loads this library holds its own copy
// until the process ends.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Packs a system's values for lane 0 of tile 0 (tile width kSysTileW): L+U position p receives the
// sum of its triplets' values in triplet order (pk_ptr / pk_idx), the other lanes of the position's
// block zero; y row k receives rhs[prow[k]] in lane 0, zero elsewhere; the tile's status marks lane 0
// running and the rest padding.  Every value the LU kernel reads is written here.
__global__ __launch_bounds__(256) void k_sys_stage(const int* __restrict__ pk_ptr, const int* __restrict__ pk_idx,
                                                   const double* __restrict__ vals, int nlu,
                                                   const int* __restrict__ prow, const double* __restrict__ rhs, int n,
                                                   double* __restrict__ lu, double* __restrict__ y,
                                                   int* __restrict__ status, int* __restrict__ iters) {
  constexpr int V = lu_block(kSysTileW) * kSysTileW;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nlu) {
    double acc = 0.0;
    for (int e = pk_ptr[i]; e < pk_ptr[i + 1]; ++e) acc += vals[pk_idx[e]];
    double* v = lu + (size_t)i * V;
    v[0] = acc;
#pragma unroll
    for (int l = 1; l < V; ++l) v[l] = 0.0;
  }
  if (i < n) {
    double* r = y + (size_t)i * kSysTileW;
    r[0] = rhs[prow[i]];
#pragma unroll
    for (int l = 1; l < kSysTileW; ++l) r[l] = 0.0;
  }
  if (i < kSysTileW) {
    status[i] = i == 0 ? kRunning : kPadLane;
    iters[i] = 0;
  }
}

// x[q[k]] = lane 0 of y row k.
__global__ __launch_bounds__(256) void k_sys_unpack(const double* __restrict__ y, const int* __restrict__ q, int n,
                                                    double* __restrict__ x) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) x[q[k]] = y[(size_t)k * kSysTileW];
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
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 8: f(std::integral_constant<int, 8>{}); break;
    case kWideTile: f(std::integral_constant<int, kWideTile>{}); break;
    default: f(std::integral_constant<int, 16>{}); break;
  }
}

template <typename T>
size_t lu_bytes(const DevTopo& t, int tile, bool factor) {
  if (t.part_a && tile != kWideTile) return lu_parts_lds_bytes(t.max_col, t.window + kRingGap + 1, tile, sizeof(T), factor);
  return lu_lds_bytes(t.max_col, t.window + kRingGap + 1, tile, sizeof(T), factor);
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
  static size_t granted_f = 65536, granted_s = 65536, granted_pf = 65536, granted_ps = 65536;
  const int cap = t.max_col < 1 ? 1 : t.max_col;
  if constexpr (TW == 8 && std::is_same<T, float>::value) {
    if (t.part_a && t.groups == kLanes / 4 && t.part_groups == kLanes / 2) {
      // quad tiles: the refactor on 2-lane groups, four scenarios per lane; the substitutions on
      // the 4-lane groups of pair tiles
      using P = pair2<T>;
      using Q = pair2<P>;
      static size_t granted_hf = 65536, granted_hs = 65536;
      const size_t shmem = lu_quad_lds_bytes(t.max_col, t.window + kRingGap + 1, 4, sizeof(P), 2, sizeof(Q), factor);
      if (factor) {
        allow_lds(k_lu_parts<P, 4, true, Q, 2>, shmem, granted_hf);
        hipLaunchKernelGGL((k_lu_parts<P, 4, true, Q, 2>), dim3(w.n_tiles), dim3(64), shmem, stream, t, w, cap, iter);
      } else {
        allow_lds(k_lu_parts<P, 4, false, Q, 2>, shmem, granted_hs);
        hipLaunchKernelGGL((k_lu_parts<P, 4, false, Q, 2>), dim3(w.n_tiles), dim3(64), shmem, stream, t, w, cap, iter);
      }
      return;
    }
  }
  if constexpr (TW == 8 && (std::is_same<T, float>::value || std::is_same<T, double>::value)) {
    if (t.part_a && t.groups == kLanes / 4 && t.part_groups == kLanes / 4) {
      // pair tiles: the schedule of 4-lane groups, two scenarios per lane
      using P = pair2<T>;
      static size_t granted_qf = 65536, granted_qs = 65536;
      const size_t shmem = lu_bytes<P>(t, 4, factor);
      if (factor) {
        allow_lds(k_lu_parts<P, 4, true>, shmem, granted_qf);
        hipLaunchKernelGGL((k_lu_parts<P, 4, true>), dim3(w.n_tiles), dim3(64), shmem, stream, t, w, cap, iter);
      } else {
        allow_lds(k_lu_parts<P, 4, false>, shmem, granted_qs);
        hipLaunchKernelGGL((k_lu_parts<P, 4, false>), dim3(w.n_tiles), dim3(64), shmem, stream, t, w, cap, iter);
      }
      return;
    }
  }
  const size_t shmem = lu_bytes<T>(t, TW, factor);
  if constexpr (TW != kWideTile) {
    if (t.part_a && t.part_groups == kLanes / TW) {
      if (factor) {
        allow_lds(k_lu_parts<T, TW, true>, shmem, granted_pf);
        hipLaunchKernelGGL((k_lu_parts<T, TW, true>), dim3(w.n_tiles), dim3(64), shmem, stream, t, w, cap, iter);
      } else {
        allow_lds(k_lu_parts<T, TW, false>, shmem, granted_ps);
        hipLaunchKernelGGL((k_lu_parts<T, TW, false>), dim3(w.n_tiles), dim3(64), shmem, stream, t, w, cap, iter);
      }
      return;
    }
  }
  if (factor) {
    allow_lds(k_lu<T, TW, true>, shmem, granted_f);
    hipLaunchKernelGGL((k_lu<T, TW, true>), dim3(w.n_tiles), dim3(64), shmem, stream, t, w, cap, iter);
  } else {
    allow_lds(k_lu<T, TW, false>, shmem, granted_s);
    hipLaunchKernelGGL((k_lu<T, TW, false>), dim3(w.n_tiles), dim3(64), shmem, stream, t, w, cap, iter);
  }
}

template <typename T, int TW>
void run_lu_queue(const DevTopo* topos, const DevWork* works, const int2* items, int n_items, int* counter,
                  size_t shmem, int grid, bool factor, hipStream_t stream) {
  static size_t granted_f = 65536, granted_s = 65536;
  hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, stream, counter, 1);
  if (factor) {
    allow_lds(k_lu_queue<T, TW, true>, shmem, granted_f);
    hipLaunchKernelGGL((k_lu_queue<T, TW, true>), dim3(grid), dim3(64), shmem, stream, topos, works, items, n_items,
                       counter);
  } else {
    allow_lds(k_lu_queue<T, TW, false>, shmem, granted_s);
    hipLaunchKernelGGL((k_lu_queue<T, TW, false>), dim3(grid), dim3(64), shmem, stream, topos, works, items, n_items,
                       counter);
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

void launch_mismatch(const DevTopo& t, const DevWork& w, bool fp32, hipStream_t stream, bool fill) {
  with_tile(w.tile, [&](auto tw) {
    constexpr int TW = decltype(tw)::value;
    dim3 grid(elem_blocks(t.n_bus, TW), w.n_tiles);
    if (fill && t.y_lupos) {
      if (fp32) hipLaunchKernelGGL((k_mismatch<float, TW, true>), grid, dim3(256), 0, stream, t, w);
      else hipLaunchKernelGGL((k_mismatch<double, TW, true>), grid, dim3(256), 0, stream, t, w);
      return;
    }
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
    const int tb = (w.n_tiles + 256 / TW - 1) / (256 / TW);
    if (!t.fb_ptr) {  // (no bus-pair recipes: one position per workgroup)
      dim3 grid(t.nnz_lu, tb);
      if (fp32) hipLaunchKernelGGL((k_fill<float, TW>), grid, dim3(256), 0, stream, t, w);
      else hipLaunchKernelGGL((k_fill<double, TW>), grid, dim3(256), 0, stream, t, w);
      return;
    }
    (void)tb;
    constexpr int B = lu_block(TW), V = B * TW;
    dim3 grid(elem_blocks(t.n_bus, V <= kLanes ? V : TW), V <= kLanes ? (w.n_tiles + B - 1) / B : w.n_tiles);
    if (fp32) hipLaunchKernelGGL((k_fill_tile<float, TW>), grid, dim3(256), 0, stream, t, w);
    else hipLaunchKernelGGL((k_fill_tile<double, TW>), grid, dim3(256), 0, stream, t, w);
    if (t.n_fz > 0 && !t.stage_zero) {  // (else the refactor stages the fill entries as zero)
      constexpr int B = lu_block(TW);
      const int fb = (t.n_fz + 255) / 256;
      dim3 zgrid(fb < 64 ? fb : 64, (w.n_tiles + B - 1) / B);
      if (fp32) hipLaunchKernelGGL((k_fill_zero<float, TW>), zgrid, dim3(256), 0, stream, t, w);
      else hipLaunchKernelGGL((k_fill_zero<double, TW>), zgrid, dim3(256), 0, stream, t, w);
    }
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

void launch_lu_queue(int tile, const DevTopo* topos, const DevWork* works, const int2* items, int n_items,
                     int* counter, size_t shmem, int grid, bool fp32, bool factor, hipStream_t stream) {
  if (n_items <= 0) return;
  with_tile(tile, [&](auto tw) {
    constexpr int TW = decltype(tw)::value;
    if (fp32) run_lu_queue<float, TW>(topos, works, items, n_items, counter, shmem, grid, factor, stream);
    else run_lu_queue<double, TW>(topos, works, items, n_items, counter, shmem, grid, factor, stream);
  });
}

void launch_iter_advance(int* iter_dev, hipStream_t stream) {
  hipLaunchKernelGGL(k_iter_advance, dim3(1), dim3(64), 0, stream, iter_dev);
}

void launch_zero_ints(int* p, int n, hipStream_t stream) {
  if (n > 0) hipLaunchKernelGGL(k_zero_ints, dim3((n + 255) / 256), dim3(256), 0, stream, p, n);
}

size_t lu_shmem_tile(const DevTopo& t, bool fp32, bool factor, int tile) {
  if (tile == 8 && t.part_a && t.groups == kLanes / 4) {  // pair tiles (and quad-tile refactors)
    if (fp32 && t.part_groups == kLanes / 2)
      return lu_quad_lds_bytes(t.max_col, t.window + kRingGap + 1, 4, sizeof(f2), 2, sizeof(f4), factor);
    return fp32 ? lu_bytes<f2>(t, 4, factor) : lu_bytes<d2>(t, 4, factor);
  }
  return fp32 ? lu_bytes<float>(t, tile, factor) : lu_bytes<double>(t, tile, factor);
}

void launch_refine(const DevTopo& t, const DevWork& w, double* dvr, double* dvi, int iter, hipStream_t stream) {
  with_tile(w.tile, [&](auto tw) {
    constexpr int TW = decltype(tw)::value;
    dim3 grid(elem_blocks(t.n_bus, TW), w.n_tiles);
    hipLaunchKernelGGL((k_dv<float, TW>), grid, dim3(256), 0, stream, t, w, dvr, dvi, iter);
    hipLaunchKernelGGL((k_resid<float, TW>), grid, dim3(256), 0, stream, t, w, dvr, dvi, iter);
  });
}

void launch_update(const DevTopo& t, const DevWork& w, bool fp32, bool mixed, int iter, hipStream_t stream) {
  with_tile(w.tile, [&](auto tw) {
    constexpr int TW = decltype(tw)::value;
    dim3 grid(elem_blocks(t.n_bus, TW), w.n_tiles);
    if (fp32) {
      if (mixed) hipLaunchKernelGGL((k_update<float, TW, true>), grid, dim3(256), 0, stream, t, w, iter);
      else hipLaunchKernelGGL((k_update<float, TW, false>), grid, dim3(256), 0, stream, t, w, iter);
    } else {
      hipLaunchKernelGGL((k_update<double, TW, false>), grid, dim3(256), 0, stream, t, w, iter);
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

void launch_warm_start(const DevTopo& t, const DevWork& w, const double* warm, const int* ok, hipStream_t stream) {
  with_tile(w.tile, [&](auto tw) {
    constexpr int TW = decltype(tw)::value;
    hipLaunchKernelGGL(k_warm_start<TW>, dim3(elem_blocks(t.n_bus, TW), w.n_tiles), dim3(256), 0, stream, t, w, warm,
                       ok);
  });
}

void launch_keep_warm(const DevTopo& t, const DevWork& w, double* warm, int* ok, hipStream_t stream) {
  with_tile(w.tile, [&](auto tw) {
    constexpr int TW = decltype(tw)::value;
    hipLaunchKernelGGL(k_keep_warm<TW>, dim3(1), dim3(256), 0, stream, t, w, warm, ok);
  });
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

hipError_t scratch_base(void** out) {
  // the array's address on each device, looked up once per device with that device current
  static std::mutex mu;
  static void* base[kMaxDevices] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!base[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    base[dev] = p;
  }
  *out = base[dev];
  return hipSuccess;
}

void launch_sys_stage(const int* pk_ptr, const int* pk_idx, const double* vals, int nlu, const int* prow,
                      const double* rhs, int n, double* lu, double* y, int* status, int* iters, hipStream_t stream) {
  const int m = std::max(std::max(nlu, n), kSysTileW);
  hipLaunchKernelGGL(k_sys_stage, dim3((m + 255) / 256), dim3(256), 0, stream, pk_ptr, pk_idx, vals, nlu, prow, rhs, n,
                     lu, y, status, iters);
}

void launch_sys_unpack(const double* y, const int* q, int n, double* x, hipStream_t stream) {
  if (n > 0) hipLaunchKernelGGL(k_sys_unpack, dim3((n + 255) / 256), dim3(256), 0, stream, y, q, n, x);
}

void launch_lane0(const double* src, double* dst, int count, bool to_tile, int stride, hipStream_t stream) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_lane0, dim3((count + 255) / 256), dim3(256), 0, stream, src, dst, count, to_tile ? 1 : 0,
                     stride);
}

}  // namespace gx
