// HIP kernels for the batched Newton-Raphson power flow on CDNA4 (gfx950).
//
// One lane = one scenario.  A tile of 64 scenarios shares every index (the Y-bus, the Jacobian
// recipes, the LU pattern), so all index arithmetic is wave-uniform (scalar loads, no divergence)
// and every value access is one coalesced 64-lane load of the same element of 64 scenarios.
//
// Kernel  <- reference function it replaces
//   k_mismatch  power_injections + mismatch (src/network.rs:680-701, src/solver.rs:681-697)
//   k_decide    convergence test and history (src/solver.rs:697-703)
//   k_fill      JacobianPattern::fill_into (src/jacobian.rs:210-250), written straight into the
//               packed L+U storage at each entry's pivot position (pack_values_slice,
//               src/klu_native/mod.rs:123-125, with the zero fill KLU's refactor scatters into)
//   k_lu        refactor_block_in_place (src/klu_native/kernel.rs:476-524) + solve
//               (src/klu_native/solve.rs:56-102): left-looking, pivots fixed, forward
//               substitution fused into the column loop, then back substitution
//   k_update    voltage update (src/solver.rs:722-732)
//   k_dv/k_resid  FP64 residual f - J*dx for the mixed-precision refinement sweep, evaluated
//               from the linearised injections dS = dV*conj(I) + V*conj(Y*dV) (the exact
//               derivative that the H/N/M/L formulas tabulate)
#include "gx_device.hpp"

#include <cmath>

namespace gx {
namespace {

__device__ __forceinline__ size_t at(int tile, int count, int e) {
  return ((size_t)tile * (size_t)count + (size_t)e) * kLanes;
}

__device__ __forceinline__ bool tile_running(const int* status, int tile) {
  const int lane = threadIdx.x & (kLanes - 1);
  return __any(status[tile * kLanes + lane] == kRunning);
}

// ---------------------------------------------------------------- layout conversion
// 64 scenarios x 64 elements per block, transposed through LDS so both sides are coalesced.
__global__ __launch_bounds__(256) void k_to_tiles(const double* __restrict__ src, double* __restrict__ dst,
                                                  int n_scen, int count) {
  __shared__ double buf[64][65];
  const int tile = blockIdx.y, e0 = blockIdx.x * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int r = ty; r < 64; r += 4) {  // r = scenario within tile
    int s = tile * 64 + r;
    s = s < n_scen ? s : n_scen - 1;  // padding lanes repeat the last scenario
    int e = e0 + tx;
    buf[r][tx] = (e < count) ? src[(size_t)s * count + e] : 0.0;
  }
  __syncthreads();
  for (int c = ty; c < 64; c += 4) {  // c = element
    int e = e0 + c;
    if (e < count) dst[at(tile, count, e) + tx] = buf[tx][c];
  }
}

__global__ __launch_bounds__(256) void k_from_tiles(const double* __restrict__ src, double* __restrict__ dst,
                                                    int n_scen, int count) {
  __shared__ double buf[64][65];
  const int tile = blockIdx.y, e0 = blockIdx.x * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int c = ty; c < 64; c += 4) {
    int e = e0 + c;
    buf[tx][c] = (e < count) ? src[at(tile, count, e) + tx] : 0.0;
  }
  __syncthreads();
  for (int r = ty; r < 64; r += 4) {
    int s = tile * 64 + r;
    int e = e0 + tx;
    if (s < n_scen && e < count) dst[(size_t)s * count + e] = buf[r][tx];
  }
}

// ---------------------------------------------------------------- per-iteration kernels
__global__ __launch_bounds__(256) void k_init(DevTopo t, DevWork w) {
  const int tile = blockIdx.y, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int s = tile * kLanes + lane;
  if (blockIdx.x == 0 && wave == 0) {
    w.status[s] = s < w.n_scen ? kRunning : kPad;
    w.iters[s] = 0;
    w.maxmis[s] = 0ull;
  }
  for (int i = blockIdx.x * 4 + wave; i < t.n_bus; i += gridDim.x * 4) {
    size_t o = at(tile, t.n_bus, i) + lane;
    double m = w.vm[o], a = w.va[o];
    double sn, cs;
    sincos(a, &sn, &cs);
    w.vr[o] = m * cs;
    w.vi[o] = m * sn;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_mismatch(DevTopo t, DevWork w) {
  const int tile = blockIdx.y, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (!tile_running(w.status, tile)) return;
  const int s = tile * kLanes + lane;
  const bool run = w.status[s] == kRunning;
  const size_t bb = at(tile, t.n_bus, 0) + lane;
  const size_t ub = at(tile, t.n, 0) + lane;
  T* __restrict__ y = reinterpret_cast<T*>(w.y);
  double mx = 0.0;
  for (int i = blockIdx.x * 4 + wave; i < t.n_bus; i += gridDim.x * 4) {
    double ir = 0.0, ii = 0.0;
    const int p1 = t.y_rowptr[i], p2 = t.y_rowptr[i + 1];
    for (int p = p1; p < p2; ++p) {
      const int j = t.y_col[p];
      const double g = t.y_g[p], b = t.y_b[p];
      const double vr = w.vr[bb + (size_t)j * kLanes], vi = w.vi[bb + (size_t)j * kLanes];
      ir += g * vr - b * vi;
      ii += g * vi + b * vr;
    }
    const size_t o = bb + (size_t)i * kLanes;
    const double vr = w.vr[o], vi = w.vi[o];
    const double pcal = vr * ir + vi * ii;
    const double qcal = vi * ir - vr * ii;
    w.pc[o] = pcal;
    w.qc[o] = qcal;
    const int rp = t.rpos_p[i];
    if (rp >= 0) {
      const double fp = w.ps[o] - pcal;
      w.f[ub + (size_t)rp * kLanes] = fp;
      y[ub + (size_t)rp * kLanes] = (T)fp;
      mx = fmax(mx, fabs(fp));
    }
    const int rq = t.rpos_q[i];
    if (rq >= 0) {
      const double fq = w.qs[o] - qcal;
      w.f[ub + (size_t)rq * kLanes] = fq;
      y[ub + (size_t)rq * kLanes] = (T)fq;
      mx = fmax(mx, fabs(fq));
    }
  }
  if (run && mx > 0.0) atomicMax(&w.maxmis[s], (unsigned long long)__double_as_longlong(mx));
}

__global__ void k_decide(DevWork w, int iter) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= w.n_tiles * kLanes) return;
  if (w.status[s] != kRunning) return;
  const double m = __longlong_as_double((long long)w.maxmis[s]);
  w.maxmis[s] = 0ull;
  w.history[(size_t)s * w.max_iter + iter] = m;
  if (m < w.tol) {
    w.status[s] = kConverged;
    w.iters[s] = iter + 1;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_fill(DevTopo t, DevWork w) {
  const int tile = blockIdx.y, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (!tile_running(w.status, tile)) return;
  const size_t bb = at(tile, t.n_bus, 0) + lane;
  T* __restrict__ lu = reinterpret_cast<T*>(w.lu) + at(tile, t.nnz_lu, 0) + lane;
  constexpr int kChunk = 32;
  for (int c = blockIdx.x * 4 + wave; c * kChunk < t.nnz_lu; c += gridDim.x * 4) {
    const int pe = min(t.nnz_lu, (c + 1) * kChunk);
    for (int p = c * kChunk; p < pe; ++p) {
      const int kind = t.lu_kind[p];
      double v = 0.0;
      if (kind != kFill) {
        const int i = t.lu_bi[p], k = t.lu_bk[p];
        const double g = t.lu_g[p], b = t.lu_b[p];
        const size_t oi = bb + (size_t)i * kLanes;
        const double vmi = w.vm[oi];
        switch (kind) {
          case kHii: v = -w.qc[oi] - vmi * vmi * b; break;
          case kNii: v = w.pc[oi] / vmi + vmi * g; break;
          case kMii: v = w.pc[oi] - vmi * vmi * g; break;
          case kLii: v = w.qc[oi] / vmi - vmi * b; break;
          default: {
            // V_i V_k cos(t_i - t_k) and V_i V_k sin(t_i - t_k) from the rectangular voltages
            const size_t ok = bb + (size_t)k * kLanes;
            const double vri = w.vr[oi], vii = w.vi[oi], vrk = w.vr[ok], vik = w.vi[ok];
            const double cc = vri * vrk + vii * vik;
            const double ss = vii * vrk - vri * vik;
            switch (kind) {
              case kHik: v = g * ss - b * cc; break;
              case kNik: v = (g * cc + b * ss) / w.vm[ok]; break;
              case kMik: v = -(g * cc + b * ss); break;
              default: v = (g * ss - b * cc) / w.vm[ok]; break;  // kLik
            }
          }
        }
      }
      lu[(size_t)p * kLanes] = (T)v;
    }
  }
}

// Refactor (optional) + forward/back substitution for one 64-scenario tile per wave.
// Column k is staged in LDS when it fits (cap entries), updated there, then written back.
template <typename T, bool kFactor>
__global__ __launch_bounds__(64) void k_lu(DevTopo t, DevWork w, int cap, int iter) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* col = reinterpret_cast<T*>(smem_raw);
  const int tile = blockIdx.x, lane = threadIdx.x;
  if (!tile_running(w.status, tile)) return;
  const int s = tile * kLanes + lane;
  T* __restrict__ lu = reinterpret_cast<T*>(w.lu) + at(tile, t.nnz_lu, 0) + lane;
  T* __restrict__ y = reinterpret_cast<T*>(w.y) + at(tile, t.n, 0) + lane;
  bool ok = true;
  const int n = t.n;
  for (int k = 0; k < n; ++k) {
    const int c0 = t.cp[k], len = t.cp[k + 1] - c0, ul = t.ulen[k];
    if (kFactor) {
      const bool staged = len <= cap;
      T* __restrict__ dst = staged ? col + lane : lu + (size_t)c0 * kLanes;
      if (staged)
        for (int e = 0; e < len; ++e) col[e * kLanes + lane] = lu[(size_t)(c0 + e) * kLanes];
      for (int u = 0; u < ul; ++u) {
        const T ujk = dst[u * kLanes];
        const int j = t.lu_row[c0 + u];
        const int l0 = t.cp[j] + t.ulen[j] + 1;
        const int ll = t.cp[j + 1] - l0;
        const uint16_t* __restrict__ off = t.upd_off + t.upd_ptr[c0 + u];
        const T* __restrict__ lsrc = lu + (size_t)l0 * kLanes;
        int e = 0;
        // targets of one update are distinct rows: batch loads before stores for ILP
        for (; e + 8 <= ll; e += 8) {
          T lv[8], xv[8];
          int o[8];
#pragma unroll
          for (int q = 0; q < 8; ++q) o[q] = off[e + q];
#pragma unroll
          for (int q = 0; q < 8; ++q) lv[q] = lsrc[(size_t)(e + q) * kLanes];
#pragma unroll
          for (int q = 0; q < 8; ++q) xv[q] = dst[o[q] * kLanes];
#pragma unroll
          for (int q = 0; q < 8; ++q) dst[o[q] * kLanes] = xv[q] - lv[q] * ujk;
        }
        for (; e < ll; ++e) {
          const int o = off[e];
          dst[o * kLanes] = dst[o * kLanes] - lsrc[(size_t)e * kLanes] * ujk;
        }
      }
      const T ukk = dst[ul * kLanes];
      if (ukk == (T)0) ok = false;
      for (int e = ul + 1; e < len; ++e) dst[e * kLanes] = dst[e * kLanes] / ukk;
      if (staged)
        for (int e = 0; e < len; ++e) lu[(size_t)(c0 + e) * kLanes] = col[e * kLanes + lane];
    }
    // forward substitution with the finished column k of L
    const T yk = y[(size_t)k * kLanes];
    for (int e = ul + 1; e < len; ++e) {
      const int i = t.lu_row[c0 + e];
      y[(size_t)i * kLanes] -= lu[(size_t)(c0 + e) * kLanes] * yk;
    }
  }
  for (int k = n - 1; k >= 0; --k) {
    const int c0 = t.cp[k], ul = t.ulen[k];
    const T xk = y[(size_t)k * kLanes] / lu[(size_t)(c0 + ul) * kLanes];
    y[(size_t)k * kLanes] = xk;
    if (!isfinite((double)xk)) ok = false;
    for (int u = 0; u < ul; ++u) {
      const int j = t.lu_row[c0 + u];
      y[(size_t)j * kLanes] -= lu[(size_t)(c0 + u) * kLanes] * xk;
    }
  }
  if (!ok && w.status[s] == kRunning) {
    w.status[s] = kSingular;
    w.iters[s] = iter + 1;
  }
}

// Mixed precision, step 1 of the refinement: dx1 (FP64 copy of the FP32 solve) and
// dV_i = V_i * (d|V|_i / |V|_i + j dtheta_i) per bus.
template <typename T>
__global__ __launch_bounds__(256) void k_dv(DevTopo t, DevWork w, double* __restrict__ dvr, double* __restrict__ dvi) {
  const int tile = blockIdx.y, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (!tile_running(w.status, tile)) return;
  const size_t bb = at(tile, t.n_bus, 0) + lane;
  const size_t ub = at(tile, t.n, 0) + lane;
  const T* __restrict__ y = reinterpret_cast<const T*>(w.y);
  for (int i = blockIdx.x * 4 + wave; i < t.n_bus; i += gridDim.x * 4) {
    const size_t o = bb + (size_t)i * kLanes;
    double da = 0.0, dm = 0.0;
    const int xa = t.xpos_a[i], xm = t.xpos_m[i];
    if (xa >= 0) {
      da = (double)y[ub + (size_t)xa * kLanes];
      w.dx[ub + (size_t)xa * kLanes] = da;
    }
    if (xm >= 0) {
      dm = (double)y[ub + (size_t)xm * kLanes];
      w.dx[ub + (size_t)xm * kLanes] = dm;
    }
    const double a = dm / w.vm[o];
    const double vr = w.vr[o], vi = w.vi[o];
    dvr[o] = vr * a - vi * da;
    dvi[o] = vr * da + vi * a;
  }
}

// Mixed precision, step 2: r = f - J*dx1 in FP64, written (in LU precision) as the next rhs.
template <typename T>
__global__ __launch_bounds__(256) void k_resid(DevTopo t, DevWork w, const double* __restrict__ dvr,
                                               const double* __restrict__ dvi) {
  const int tile = blockIdx.y, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (!tile_running(w.status, tile)) return;
  const size_t bb = at(tile, t.n_bus, 0) + lane;
  const size_t ub = at(tile, t.n, 0) + lane;
  T* __restrict__ y = reinterpret_cast<T*>(w.y);
  for (int i = blockIdx.x * 4 + wave; i < t.n_bus; i += gridDim.x * 4) {
    const int rp = t.rpos_p[i], rq = t.rpos_q[i];
    if (rp < 0 && rq < 0) continue;
    double ar = 0.0, ai = 0.0;
    for (int p = t.y_rowptr[i]; p < t.y_rowptr[i + 1]; ++p) {
      const int j = t.y_col[p];
      const double g = t.y_g[p], b = t.y_b[p];
      const double dr = dvr[bb + (size_t)j * kLanes], di = dvi[bb + (size_t)j * kLanes];
      ar += g * dr - b * di;
      ai += g * di + b * dr;
    }
    const size_t o = bb + (size_t)i * kLanes;
    const double vr = w.vr[o], vi = w.vi[o];
    // dV_i / V_i = a + j b
    const int xa = t.xpos_a[i], xm = t.xpos_m[i];
    const double bq = xa >= 0 ? w.dx[ub + (size_t)xa * kLanes] : 0.0;
    const double aq = xm >= 0 ? w.dx[ub + (size_t)xm * kLanes] / w.vm[o] : 0.0;
    const double pc = w.pc[o], qc = w.qc[o];
    const double dsr = (aq * pc - bq * qc) + (vr * ar + vi * ai);
    const double dsi = (aq * qc + bq * pc) + (vi * ar - vr * ai);
    if (rp >= 0) y[ub + (size_t)rp * kLanes] = (T)(w.f[ub + (size_t)rp * kLanes] - dsr);
    if (rq >= 0) y[ub + (size_t)rq * kLanes] = (T)(w.f[ub + (size_t)rq * kLanes] - dsi);
  }
}

template <typename T, bool kMixed>
__global__ __launch_bounds__(256) void k_update(DevTopo t, DevWork w) {
  const int tile = blockIdx.y, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (!tile_running(w.status, tile)) return;
  const int s = tile * kLanes + lane;
  if (w.status[s] != kRunning) return;  // lanes diverge only here; no barriers follow
  const size_t bb = at(tile, t.n_bus, 0) + lane;
  const size_t ub = at(tile, t.n, 0) + lane;
  const T* __restrict__ y = reinterpret_cast<const T*>(w.y);
  for (int i = blockIdx.x * 4 + wave; i < t.n_bus; i += gridDim.x * 4) {
    const int xa = t.xpos_a[i], xm = t.xpos_m[i];
    if (xa < 0 && xm < 0) continue;
    const size_t o = bb + (size_t)i * kLanes;
    double a = w.va[o], m = w.vm[o];
    if (xa >= 0) {
      double d = (double)y[ub + (size_t)xa * kLanes];
      if (kMixed) d += w.dx[ub + (size_t)xa * kLanes];
      a += d;
    }
    if (xm >= 0) {
      double d = (double)y[ub + (size_t)xm * kLanes];
      if (kMixed) d += w.dx[ub + (size_t)xm * kLanes];
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

__global__ void k_finish(DevWork w) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= w.n_tiles * kLanes) return;
  if (w.status[s] == kRunning) {
    w.status[s] = kMaxIter;
    w.iters[s] = w.max_iter;
  }
}

// Copies a compact vector into / out of lane 0 of an interleaved single-tile array.
__global__ void k_lane0(const double* __restrict__ src, double* __restrict__ dst, int count, int to_tile) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= count) return;
  if (to_tile) dst[(size_t)e * kLanes] = src[e];
  else dst[e] = src[(size_t)e * kLanes];
}

inline int bus_blocks(int n_bus) {
  int b = (n_bus + 4 * 8 - 1) / (4 * 8);  // about 8 buses per wave
  return b < 1 ? 1 : (b > 1024 ? 1024 : b);
}

}  // namespace

void launch_to_tiles(const double* src, double* dst, int n_scen, int n_tiles, int count, hipStream_t stream) {
  if (count <= 0 || n_tiles <= 0) return;
  dim3 grid((count + 63) / 64, n_tiles);
  hipLaunchKernelGGL(k_to_tiles, grid, dim3(256), 0, stream, src, dst, n_scen, count);
}

void launch_from_tiles(const double* src, double* dst, int n_scen, int count, hipStream_t stream) {
  int n_tiles = (n_scen + kLanes - 1) / kLanes;
  if (count <= 0 || n_tiles <= 0) return;
  dim3 grid((count + 63) / 64, n_tiles);
  hipLaunchKernelGGL(k_from_tiles, grid, dim3(256), 0, stream, src, dst, n_scen, count);
}

void launch_init(const DevTopo& t, const DevWork& w, hipStream_t stream) {
  dim3 grid(bus_blocks(t.n_bus), w.n_tiles);
  hipLaunchKernelGGL(k_init, grid, dim3(256), 0, stream, t, w);
}

void launch_mismatch(const DevTopo& t, const DevWork& w, bool fp32, hipStream_t stream);

void launch_decide(const DevTopo& t, const DevWork& w, int iter, hipStream_t stream) {
  (void)t;
  int threads = w.n_tiles * kLanes;
  hipLaunchKernelGGL(k_decide, dim3((threads + 255) / 256), dim3(256), 0, stream, w, iter);
}

void launch_fill(const DevTopo& t, const DevWork& w, bool fp32, hipStream_t stream) {
  int chunks = (t.nnz_lu + 31) / 32;
  int bx = (chunks + 3) / 4;
  bx = bx < 1 ? 1 : (bx > 2048 ? 2048 : bx);
  dim3 grid(bx, w.n_tiles);
  if (fp32) hipLaunchKernelGGL(k_fill<float>, grid, dim3(256), 0, stream, t, w);
  else hipLaunchKernelGGL(k_fill<double>, grid, dim3(256), 0, stream, t, w);
}

void launch_lu(const DevTopo& t, const DevWork& w, bool fp32, bool factor, int iter, hipStream_t stream) {
  const size_t es = fp32 ? sizeof(float) : sizeof(double);
  int cap = (int)(49152 / (kLanes * es));
  if (cap > t.max_col) cap = t.max_col;
  const size_t shmem = factor ? (size_t)cap * kLanes * es : 0;
  dim3 grid(w.n_tiles);
  if (fp32) {
    if (factor) hipLaunchKernelGGL((k_lu<float, true>), grid, dim3(64), shmem, stream, t, w, cap, iter);
    else hipLaunchKernelGGL((k_lu<float, false>), grid, dim3(64), shmem, stream, t, w, cap, iter);
  } else {
    if (factor) hipLaunchKernelGGL((k_lu<double, true>), grid, dim3(64), shmem, stream, t, w, cap, iter);
    else hipLaunchKernelGGL((k_lu<double, false>), grid, dim3(64), shmem, stream, t, w, cap, iter);
  }
}

void launch_mismatch(const DevTopo& t, const DevWork& w, bool fp32, hipStream_t stream) {
  dim3 grid(bus_blocks(t.n_bus), w.n_tiles);
  if (fp32) hipLaunchKernelGGL(k_mismatch<float>, grid, dim3(256), 0, stream, t, w);
  else hipLaunchKernelGGL(k_mismatch<double>, grid, dim3(256), 0, stream, t, w);
}

void launch_refine(const DevTopo& t, const DevWork& w, double* dvr, double* dvi, hipStream_t stream) {
  dim3 grid(bus_blocks(t.n_bus), w.n_tiles);
  hipLaunchKernelGGL(k_dv<float>, grid, dim3(256), 0, stream, t, w, dvr, dvi);
  hipLaunchKernelGGL(k_resid<float>, grid, dim3(256), 0, stream, t, w, dvr, dvi);
}

void launch_update(const DevTopo& t, const DevWork& w, bool fp32, bool mixed, hipStream_t stream) {
  dim3 grid(bus_blocks(t.n_bus), w.n_tiles);
  if (fp32) {
    if (mixed) hipLaunchKernelGGL((k_update<float, true>), grid, dim3(256), 0, stream, t, w);
    else hipLaunchKernelGGL((k_update<float, false>), grid, dim3(256), 0, stream, t, w);
  } else {
    hipLaunchKernelGGL((k_update<double, false>), grid, dim3(256), 0, stream, t, w);
  }
}

void launch_lane0(const double* src, double* dst, int count, bool to_tile, hipStream_t stream) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_lane0, dim3((count + 255) / 256), dim3(256), 0, stream, src, dst, count, to_tile ? 1 : 0);
}

void launch_finish(const DevWork& w, hipStream_t stream) {
  int threads = w.n_tiles * kLanes;
  hipLaunchKernelGGL(k_finish, dim3((threads + 255) / 256), dim3(256), 0, stream, w);
}

}  // namespace gx
