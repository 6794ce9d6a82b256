// Host side of gridox_hip: Y-bus build, Jacobian pattern, symbolic phase, device upload and the
// batched Newton-Raphson driver behind the C ABI in include/gridox_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <queue>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "gridox_hip.h"
#include "gx_device.hpp"
#include "gx_symbolic.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define GX_HIP(expr)                                                                 \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) return fail(GX_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Device allocation owned by a handle.
struct DevMem {
  void* p = nullptr;
  size_t bytes = 0;
  DevMem() = default;
  DevMem(const DevMem&) = delete;
  DevMem& operator=(const DevMem&) = delete;
  ~DevMem() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  int reserve(size_t b) {
    if (b <= bytes && p) return GX_OK;
    release();
    if (b == 0) b = 16;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(GX_ERR_DEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    bytes = b;
    return GX_OK;
  }
  template <class T>
  int upload(const std::vector<T>& v) {
    int rc = reserve(sizeof(T) * v.size());
    if (rc) return rc;
    if (!v.empty()) GX_HIP(hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return GX_OK;
  }
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

struct cx {
  double re, im;
};
inline cx cadd(cx a, cx b) { return {a.re + b.re, a.im + b.im}; }
inline cx cmul(cx a, cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline cx cdiv(cx a, cx b) {
  double ns = b.re * b.re + b.im * b.im;
  return {(a.re * b.re + a.im * b.im) / ns, (a.im * b.re - a.re * b.im) / ns};
}
inline cx cneg(cx a) { return {-a.re, -a.im}; }
inline cx cconj(cx a) { return {a.re, -a.im}; }

// Packed L+U schedule derived from an LuPattern (shared by the topology and the plain system).
struct LuSchedule {
  std::vector<int> upd_ptr;
  std::vector<uint16_t> upd_off;
  int max_col = 0;
  long long n_upd = 0;
};

int build_schedule(const gx::LuPattern& lu, LuSchedule& s) {
  const int n = lu.n;
  const int nnz = lu.nnz();
  s.max_col = 0;
  for (int k = 0; k < n; ++k) s.max_col = std::max(s.max_col, lu.cp[k + 1] - lu.cp[k]);
  if (s.max_col >= 65536) return fail(GX_ERR_UNSUPPORTED, "an L+U column has 65536 or more entries");
  s.upd_ptr.assign(nnz + 1, 0);
  s.upd_off.clear();
  std::vector<int> where(n, -1);
  for (int k = 0; k < n; ++k) {
    const int c0 = lu.cp[k], c1 = lu.cp[k + 1];
    for (int p = c0; p < c1; ++p) where[lu.row[p]] = p - c0;
    for (int p = c0; p < c1; ++p) {
      s.upd_ptr[p] = (int)s.upd_off.size();
      if (p - c0 < lu.ulen[k]) {
        const int j = lu.row[p];
        for (int q = lu.cp[j] + lu.ulen[j] + 1; q < lu.cp[j + 1]; ++q) {
          int o = where[lu.row[q]];
          if (o < 0) return fail(GX_ERR_ARG, "L+U pattern is not closed under elimination");
          s.upd_off.push_back((uint16_t)o);
        }
      }
    }
    for (int p = c0; p < c1; ++p) where[lu.row[p]] = -1;
  }
  s.upd_ptr[nnz] = (int)s.upd_off.size();
  if (s.upd_off.size() >= (size_t)0x7fffffff) return fail(GX_ERR_UNSUPPORTED, "update schedule too large");
  s.n_upd = (long long)s.upd_off.size();
  return GX_OK;
}

// Refactor update rounds for G lane groups (the layout is described with DevTopo::slots).
// Column k's update items (target row offset t, source U offset u, L(i,j) at position p) are taken
// in the order of the sequential left-looking refactor (j ascending, then i ascending); each goes to
// the earliest round after every earlier item that targets t or u (so every entry receives its
// updates in that same order, and U(j,k) is final before it is used), into any free slot (the
// lanes of one wavefront see each other's global stores in program order); no other item of the
// round may target t or read t.
struct RoundPlan {
  std::vector<int> chunk0, rounds;
  std::vector<int> slots;  // pairs {x, word}
};

int build_rounds(const gx::LuPattern& lu, const LuSchedule& s, int G, RoundPlan& out) {
  const int n = lu.n, NB = gx::kChunkRounds;
  if (s.max_col >= 32768) return fail(GX_ERR_UNSUPPORTED, "an L+U column has 32768 or more entries");
  out.chunk0.assign(n + 1, 0);
  out.rounds.assign(n, 0);
  out.slots.clear();
  std::vector<int> last;
  std::vector<int> sx, sw;                  // per round * G slot
  std::vector<std::vector<int>> rt, rs;     // per round: targets, sources
  for (int k = 0; k < n; ++k) {
    const int c0 = lu.cp[k], len = lu.cp[k + 1] - c0, ul = lu.ulen[k];
    const int noop_w = len;  // empty slot: target the unused row just past the column
    last.assign(len, -1);
    sx.clear();
    sw.clear();
    int R = 0;
    auto add_round = [&]() {
      sx.insert(sx.end(), G, -1);
      sw.insert(sw.end(), G, noop_w);
      if ((int)rt.size() <= R) {
        rt.emplace_back();
        rs.emplace_back();
      }
      rt[R].clear();
      rs[R].clear();
      ++R;
    };
    for (int u = 0; u < ul; ++u) {
      const int j = lu.row[c0 + u], l0 = lu.cp[j] + lu.ulen[j] + 1, ll = lu.cp[j + 1] - l0;
      const int q0 = s.upd_ptr[c0 + u];
      const bool prev = j == k - 1;
      for (int e = 0; e < ll; ++e) {
        const int p = l0 + e, tg = s.upd_off[q0 + e];
        for (int r = std::max(last[tg], last[u]) + 1;; ++r) {
          while (r >= R) add_round();
          int slot = -1;
          for (int gg = 0; gg < G && slot < 0; ++gg)
            if (sx[(size_t)r * G + gg] == -1) slot = gg;
          if (slot < 0) continue;
          if (std::find(rt[r].begin(), rt[r].end(), tg) != rt[r].end()) continue;
          if (std::find(rs[r].begin(), rs[r].end(), tg) != rs[r].end()) continue;
          sx[(size_t)r * G + slot] = prev ? p - lu.cp[j] : p;
          sw[(size_t)r * G + slot] = tg | (u << 16) | (prev ? (int)0x80000000u : 0);
          rt[r].push_back(tg);
          rs[r].push_back(u);
          last[tg] = r;
          break;
        }
      }
    }
    const int nch = std::max(1, (R + NB - 1) / NB);
    out.rounds[k] = R;
    out.chunk0[k + 1] = out.chunk0[k] + nch;
    for (int r = 0; r < nch * NB; ++r)
      for (int gg = 0; gg < G; ++gg) {
        const bool real = r < R && sx[(size_t)r * G + gg] >= 0;
        out.slots.push_back(real ? sx[(size_t)r * G + gg] : 0);
        out.slots.push_back(real ? sw[(size_t)r * G + gg] : noop_w);
      }
    if (out.slots.size() >= (size_t)0x7fffffff) return fail(GX_ERR_UNSUPPORTED, "refactor round schedule too large");
  }
  // two chunks past the last one, read by the kernel's prefetch and never executed
  out.slots.insert(out.slots.end(), 2 * (size_t)2 * NB * G, 0);
  return GX_OK;
}

// Refactor by parts (narrow tiles): the columns of one elimination level depend only on columns of
// earlier levels, so a level's columns are factored together.  Column k's level is one more than the
// highest level among the j of its U column (0 if none).  Each level's columns, in ascending k, are
// cut into parts of at most `pcap` entries; parts run level by level.  A part's columns lie one after
// the other in its LDS buffer (column k at base_k), and its update items (x_t -= L(i,j) * x_u for
// column k's U entries u in ascending order, the sequential order, so the factors are bit-identical
// to the column-by-column refactor) are scheduled into joint rounds of kJointRounds * G items (the
// kernel reads the operands of a joint round's kJointRounds rounds of G items before it writes any):
// distinct targets, no target that is also read as a source, a target's updates in order and after
// its source is final.  The entries of a part past the first kPartRegs per lane group start at zero
// (the kernel clears their rows); those that are not structural fill (kind[p] == kFill when kind is
// given) are staged by the first joint rounds with items x_o = x_o - value * x_m, m = pcap + 1 the
// LDS row that holds -1 (ahead of every update of the part), so their values flow through the same
// load stream as the L values.  A registered entry whose pent_pos carries kStageZero is staged as
// zero.  Word of an item: target (bits 0-9) | source (bits 10-19) | LDS offset of an L value of the
// previous part (bits 20-29, with bit 31: read there, not from x); x: the packed position of the
// value to load (0 when none is needed).  Every part spans at least
// kPartMinChunks chunks: the kernel loads a chunk's values kPartMinChunks chunks ahead, so the values
// of a part are loaded no earlier than during the part before it, after the part two back was written.
struct PartPlan {
  std::vector<int> part_a;    // 4 per part: first entry (into pent_*), entries, first chunk, rounds
  std::vector<int> pent_pos;  // per part entry (in buffer order): its packed L+U position
  std::vector<int> pent_div;  // -1 U entry, -2 pivot, else the buffer offset of its column's pivot
  std::vector<int> slots;     // pairs {x, word} per round and lane group, chunks of part_chunk_rounds(G)
  int n_parts = 0;
  int cap = 0;                // entries of the largest part
};

constexpr bool kPartnerVisits = true;  // split visits of partner columns (below)
constexpr long kPartnerMinShared = 64;
constexpr int kPartnerMax = 1;  // partner columns per whole column (2 and 3 measured the same)

int build_parts(const gx::LuPattern& lu, const LuSchedule& s, int G, int pcap, const std::vector<int>* kind,
                PartPlan& out) {
  const int n = lu.n, NB = gx::part_chunk_rounds(G), J = gx::kJointRounds * G;
  static_assert(gx::part_chunk_rounds(32) % gx::kJointRounds == 0 && gx::part_chunk_rounds(16) % gx::kJointRounds == 0,
                "a chunk holds whole joint rounds");
  out = PartPlan();
  // (10-bit offsets in the item words; a larger part falls back to the column-by-column kernel)
  if (s.max_col > pcap || pcap >= gx::kPartMaxEntries) return fail(GX_ERR_UNSUPPORTED, "a column does not fit a refactor part");
  const int mone = pcap + 1;  // the LDS row holding -1
  std::vector<int> lev(n, 0);
  int nlev = 0;
  for (int k = 0; k < n; ++k) {
    int l = 0;
    for (int q = lu.cp[k]; q < lu.cp[k] + lu.ulen[k]; ++q) l = std::max(l, lev[lu.row[q]] + 1);
    lev[k] = l;
    nlev = std::max(nlev, l + 1);
  }
  std::vector<std::vector<int>> bucket(nlev);
  for (int k = 0; k < n; ++k) bucket[lev[k]].push_back(k);
  // Partner visits: a column c = k + 1 (up to k + kPartnerMax) at a later level than k takes the
  // updates from its leading U entries whose columns lie at levels below k's (final before k's part
  // runs) in k's part (a partial visit), each next to k's update from the same L value where k has
  // that U entry too,
  // so the second load finds the value in the cache; its entries are written back undivided.  Its
  // own part stages them again -- from the previous part's buffer when that is k's part (values
  // loaded ahead of the part would predate the write-back), else from memory -- and applies the
  // remaining updates and the division.  Every entry keeps the sequential order of its updates.
  // Only pairs whose shared sources carry kPartnerMinShared update products or more are split
  // (the others would add part entries for little), and only on 32 lane groups, whose rounds the
  // partner's items fill (10,000-bus grid: 110,200 -> 96,607 rounds); on 16 the rounds are full
  // already, and the extra parts and staging items cost more than the cached loads save.
  std::vector<char> split(n, 0);
  std::vector<int> pre_u(n, 0);  // U entries of a split column taken by its partial visit
  if (kPartnerVisits && G >= 32)
    for (int k = 0; k < n; ++k) {
      if (split[k]) continue;  // (a partner of the column before: k's part is that one's)
      const int* kb = lu.row.data() + lu.cp[k];
      const int* ke = kb + lu.ulen[k];
      int size = lu.cp[k + 1] - lu.cp[k];
      for (int c = k + 1; c < n && c <= k + kPartnerMax; ++c) {
        const int ul = lu.ulen[c], clen = lu.cp[c + 1] - lu.cp[c];
        const int* cb = lu.row.data() + lu.cp[c];
        if (lev[c] <= lev[k] || size + clen > pcap) break;  // (c must be finished after k's part)
        int m = 0;
        long shared = 0;
        for (; m < ul && lev[cb[m]] < lev[k]; ++m)
          if (std::binary_search(kb, ke, cb[m])) shared += lu.cp[cb[m] + 1] - (lu.cp[cb[m]] + lu.ulen[cb[m]] + 1);
        if (m == 0 || shared < kPartnerMinShared) break;
        split[c] = 1;
        pre_u[c] = m;
        size += clen;
      }
    }
  struct Visit {
    int col, base, u0, u1, mode;  // U entries [u0, u1) of column col; mode 0 whole, 1 partial, 2 rest
  };
  std::vector<int> part_of(n, -1), base(n, 0), pre_part(n, -1), pre_base(n, 0);
  std::vector<std::vector<Visit>> parts;
  for (int l = 0; l < nlev; ++l) {
    int fill = pcap + 1;
    for (int k : bucket[l]) {
      const int len = lu.cp[k + 1] - lu.cp[k], ul = lu.ulen[k];
      // k's partners: the split columns right after it (a whole column's; they fit one part)
      int need = len, np = 0;
      if (!split[k])
        for (int c = k + 1; c < n && split[c]; ++c, ++np) need += lu.cp[c + 1] - lu.cp[c];
      if (fill + need > pcap) {
        parts.emplace_back();
        fill = 0;
      }
      const int pi = (int)parts.size() - 1;
      parts.back().push_back({k, fill, split[k] ? pre_u[k] : 0, ul, split[k] ? 2 : 0});
      part_of[k] = pi;
      base[k] = fill;
      fill += len;
      for (int c = k + 1; c <= k + np; ++c) {
        parts.back().push_back({c, fill, 0, pre_u[c], 1});
        pre_part[c] = pi;
        pre_base[c] = fill;
        fill += lu.cp[c + 1] - lu.cp[c];
      }
    }
  }
  out.n_parts = (int)parts.size();
  out.part_a.assign(4 * (size_t)out.n_parts, 0);
  std::vector<int> last;
  std::vector<int> sx, sw;
  std::vector<std::vector<int>> rt, rs;
  int chunk = 0;
  for (int pi = 0; pi < out.n_parts; ++pi) {
    const std::vector<Visit>& vis = parts[pi];
    const int pbase = (int)out.pent_pos.size();
    for (const Visit& v : vis) {
      const int k = v.col, c0 = lu.cp[k], len = lu.cp[k + 1] - c0, ul = lu.ulen[k];
      // the rest of a split column restaged from the previous part's buffer: no value from memory
      const bool from_prev = v.mode == 2 && pre_part[k] == pi - 1;
      for (int o = 0; o < len; ++o) {
        const bool zero = v.mode == 2 ? from_prev : kind && (*kind)[c0 + o] == gx::kFill;
        out.pent_pos.push_back((c0 + o) | (zero ? gx::kStageZero : 0));
        out.pent_div.push_back(v.mode == 1 || o < ul ? -1 : (o == ul ? -2 : v.base + ul));
      }
    }
    const int Lp = (int)out.pent_pos.size() - pbase;
    out.cap = std::max(out.cap, Lp);
    const int noop_w = Lp | (Lp << 10);  // empty slot: the unused row just past the part
    last.assign(Lp + 1, -1);
    sx.clear();
    sw.clear();
    int R = 0;  // joint rounds
    auto add_round = [&]() {
      sx.insert(sx.end(), J, -1);
      sw.insert(sw.end(), J, noop_w);
      if ((int)rt.size() <= R) {
        rt.emplace_back();
        rs.emplace_back();
      }
      rt[R].clear();
      rs[R].clear();
      ++R;
    };
    // staging items, J per joint round from the first; x_o = 0 - value * (-1): entries past the
    // registered ones that are not structural fill (the kernel clears those rows), then the
    // entries restaged from the previous part's buffer
    int staged = 0;
    auto stage = [&](int o, int x, int w) {
      const int r = staged / J, slot = staged % J;
      ++staged;
      while (r >= R) add_round();
      sx[(size_t)r * J + slot] = x;
      sw[(size_t)r * J + slot] = w;
      rt[r].push_back(o);
      rs[r].push_back(mone);
      last[o] = r;
    };
    for (int o = G * gx::kPartRegs; o < Lp; ++o) {
      const int ps = out.pent_pos[pbase + o];
      if (!(ps & gx::kStageZero)) stage(o, ps, o | (mone << 10));
    }
    for (const Visit& v : vis)
      if (v.mode == 2 && pre_part[v.col] == pi - 1)
        for (int o = 0, len = lu.cp[v.col + 1] - lu.cp[v.col]; o < len; ++o)
          stage(v.base + o, 0, (v.base + o) | (mone << 10) | (int)(0x80000000u | ((unsigned)(pre_base[v.col] + o) << 20)));
    // update items, each into the earliest round after its target's previous update and its
    // source's last one where no other item targets or reads its target
    auto place = [&](int tg, int su, int x, int w) {
      for (int r = std::max(last[tg], last[su]) + 1;; ++r) {
        while (r >= R) add_round();
        int slot = -1;
        for (int gg = 0; gg < J && slot < 0; ++gg)
          if (sx[(size_t)r * J + gg] == -1) slot = gg;
        if (slot < 0) continue;
        if (std::find(rt[r].begin(), rt[r].end(), tg) != rt[r].end()) continue;
        if (std::find(rs[r].begin(), rs[r].end(), tg) != rs[r].end()) continue;
        sx[(size_t)r * J + slot] = x;
        sw[(size_t)r * J + slot] = w;
        rt[r].push_back(tg);
        rs[r].push_back(su);
        last[tg] = r;
        return;
      }
    };
    // the update of visit v's target e from its U entry u (L(i, j) x_u, j that entry's row)
    auto emit = [&](const Visit& v, int u, int e) {
      const int c0 = lu.cp[v.col], j = lu.row[c0 + u], p = lu.cp[j] + lu.ulen[j] + 1 + e;
      const int tg = v.base + s.upd_off[s.upd_ptr[c0 + u] + e], su = v.base + u;
      const bool prev = part_of[j] == pi - 1;
      place(tg, su, prev ? 0 : p,
            tg | (su << 10) | (prev ? (int)(0x80000000u | ((unsigned)(base[j] + (p - lu.cp[j])) << 20)) : 0));
    };
    for (size_t vi = 0; vi < vis.size(); ++vi) {
      const Visit& v = vis[vi];
      if (v.mode == 1) continue;  // emitted with the whole visit before them, their partner
      int np = 0;
      while (vi + 1 + np < vis.size() && vis[vi + 1 + np].mode == 1) ++np;
      if (np > kPartnerMax) return fail(GX_ERR_UNSUPPORTED, "too many partner visits in a part");
      const int c0 = lu.cp[v.col];
      auto lcount = [&](int j) { return lu.cp[j + 1] - (lu.cp[j] + lu.ulen[j] + 1); };
      int pu[kPartnerMax] = {};
      bool both[kPartnerMax] = {};
      // partner i's updates from its U entries in rows below `below`, in order
      auto partner_to = [&](int i, int below) {
        const Visit& w = vis[vi + 1 + i];
        const int* prow = lu.row.data() + lu.cp[w.col];
        for (; pu[i] < w.u1 && prow[pu[i]] < below; ++pu[i])
          for (int e = 0, ll = lcount(prow[pu[i]]); e < ll; ++e) emit(w, pu[i], e);
      };
      for (int u = v.u0; u < v.u1; ++u) {
        const int j = lu.row[c0 + u], ll = lcount(j);
        for (int i = 0; i < np; ++i) {
          partner_to(i, j);
          const Visit& w = vis[vi + 1 + i];
          both[i] = pu[i] < w.u1 && lu.row[lu.cp[w.col] + pu[i]] == j;
        }
        for (int e = 0; e < ll; ++e) {
          emit(v, u, e);
          for (int i = 0; i < np; ++i)
            if (both[i]) emit(vis[vi + 1 + i], pu[i], e);
        }
        for (int i = 0; i < np; ++i) pu[i] += both[i];
      }
      for (int i = 0; i < np; ++i) partner_to(i, n);
    }
    R *= gx::kJointRounds;  // rounds of G items: joint round r is rounds kJointRounds r + i
    const int nch = std::max(gx::kPartMinChunks, (R + NB - 1) / NB);
    out.part_a[4 * pi] = pbase;
    out.part_a[4 * pi + 1] = Lp;
    out.part_a[4 * pi + 2] = chunk;
    out.part_a[4 * pi + 3] = R;
    chunk += nch;
    for (int r = 0; r < nch * NB; ++r)
      for (int gg = 0; gg < G; ++gg) {
        const size_t at = (size_t)r * G + gg;  // slot (r % kJointRounds) G + gg of joint round r / kJointRounds
        const bool real = r < R && sx[at] >= 0;
        out.slots.push_back(real ? sx[at] : 0);
        out.slots.push_back(real ? sw[at] : noop_w);
      }
    if (out.slots.size() >= (size_t)0x7fffffff) return fail(GX_ERR_UNSUPPORTED, "refactor round schedule too large");
  }
  // chunks past the last one, read by the kernel's prefetch and never executed
  out.slots.insert(out.slots.end(), 2 * (size_t)gx::kPartAhead * NB * G, 0);
  return GX_OK;
}

// The entries the triangular solves gather instead of pushing, for an LDS window of W rows:
// L(r, j) with r - j > W and U(r, j) with j - r > W, grouped by row r (ascending j).
struct FarLists {
  std::vector<int> fptr, fpos, fcol, bptr, bpos, bcol;
};

void build_far(const gx::LuPattern& lu, int W, FarLists& f) {
  const int n = lu.n;
  f.fptr.assign(n + 1, 0);
  f.bptr.assign(n + 1, 0);
  for (int j = 0; j < n; ++j) {
    const int c0 = lu.cp[j], d = c0 + lu.ulen[j];
    for (int p = c0; p < d; ++p)
      if (j - lu.row[p] > W) f.bptr[lu.row[p] + 1]++;
    for (int p = d + 1; p < lu.cp[j + 1]; ++p)
      if (lu.row[p] - j > W) f.fptr[lu.row[p] + 1]++;
  }
  for (int r = 0; r < n; ++r) {
    f.fptr[r + 1] += f.fptr[r];
    f.bptr[r + 1] += f.bptr[r];
  }
  f.fpos.assign(f.fptr[n], 0);
  f.fcol.assign(f.fptr[n], 0);
  f.bpos.assign(f.bptr[n], 0);
  f.bcol.assign(f.bptr[n], 0);
  std::vector<int> fn(f.fptr.begin(), f.fptr.end() - 1), bn(f.bptr.begin(), f.bptr.end() - 1);
  for (int j = 0; j < n; ++j) {
    const int c0 = lu.cp[j], d = c0 + lu.ulen[j];
    for (int p = c0; p < d; ++p)
      if (j - lu.row[p] > W) {
        const int q = bn[lu.row[p]]++;
        f.bpos[q] = p;
        f.bcol[q] = j;
      }
    for (int p = d + 1; p < lu.cp[j + 1]; ++p)
      if (lu.row[p] - j > W) {
        const int q = fn[lu.row[p]]++;
        f.fpos[q] = p;
        f.fcol[q] = j;
      }
  }
}

// Update rounds of the wide tile (the layout is described with DevTopo::uitems): items taken in
// the order of the sequential left-looking refactor, each into the earliest round after every
// earlier item that targets its target or its source, where no other item targets or reads its
// target; every entry therefore receives its updates in the sequential order.
int build_uniform_rounds(const gx::LuPattern& lu, const LuSchedule& s, std::vector<int>& ur,
                         std::vector<int>& items) {
  const int n = lu.n, IPR = gx::kRoundItems;
  if (s.max_col >= 32768) return fail(GX_ERR_UNSUPPORTED, "an L+U column has 32768 or more entries");
  ur.assign(n + 1, 0);
  items.clear();
  std::vector<int> last, rx, rw, cnt;
  std::vector<std::vector<int>> rt, rs;
  for (int k = 0; k < n; ++k) {
    const int c0 = lu.cp[k], len = lu.cp[k + 1] - c0, ul = lu.ulen[k];
    last.assign(len, -1);
    rx.clear();
    rw.clear();
    cnt.clear();
    int R = 0;
    for (int u = 0; u < ul; ++u) {
      const int j = lu.row[c0 + u], l0 = lu.cp[j] + lu.ulen[j] + 1, ll = lu.cp[j + 1] - l0;
      const int q0 = s.upd_ptr[c0 + u];
      // L(:,k-1) is written back while the first round of column k is being fetched: its items
      // start at round 1
      const int r_min = j == k - 1 ? 1 : 0;
      for (int e = 0; e < ll; ++e) {
        const int tg = s.upd_off[q0 + e];
        for (int r = std::max(std::max(last[tg], last[u]) + 1, r_min);; ++r) {
          while (r >= R) {
            rx.insert(rx.end(), IPR, -1);
            rw.insert(rw.end(), IPR, len);
            cnt.push_back(0);
            if ((int)rt.size() <= R) {
              rt.emplace_back();
              rs.emplace_back();
            }
            rt[R].clear();
            rs[R].clear();
            ++R;
          }
          if (cnt[r] == IPR) continue;
          if (std::find(rt[r].begin(), rt[r].end(), tg) != rt[r].end()) continue;
          if (std::find(rs[r].begin(), rs[r].end(), tg) != rs[r].end()) continue;
          rx[(size_t)r * IPR + cnt[r]] = l0 + e;
          rw[(size_t)r * IPR + cnt[r]] = tg | (u << 16);
          ++cnt[r];
          rt[r].push_back(tg);
          rs[r].push_back(u);
          last[tg] = r;
          break;
        }
      }
    }
    if (R == 0) {  // every column owns at least one round (the kernel fetches rounds as one stream)
      rx.insert(rx.end(), IPR, -1);
      rw.insert(rw.end(), IPR, len);
      R = 1;
    }
    ur[k + 1] = ur[k] + R;
    for (size_t q = 0; q < rx.size(); ++q) {
      items.push_back(rw[q]);
      items.push_back(rx[q]);
    }
    if (items.size() >= (size_t)0x7fffffff) return fail(GX_ERR_UNSUPPORTED, "refactor round schedule too large");
  }
  return GX_OK;
}

// Device copy of a pattern's numeric schedule, pointed to by the schedule fields of a DevTopo.
// Triangular solves by steps for narrow tiles (DevTopo::back_items / fwd_items).  Back substitution:
// x_k = (y_k - sum_c U(k,c) x_c) / U(k,k); forward substitution (L unit lower): y_k = f_k - sum_j
// L(k,j) y_j.  Row k's level is one more than the highest level among the rows its entries read (0
// if none).  A row's entries, ordered by the level of the row they read (stable, so the reference
// order within a level), are cut into pieces of at most kSolvePiece entries, at most kSolvePieces
// pieces (contiguous runs, the last one the row's own): every piece but the last sums its entries
// into a partial (START from zero, ENTRY..., END into a ring row only); the last piece starts from
// y_k, takes its entries, adds the partials in piece order (COMBINE) and ends the row (END: x_k
// into y and a ring row, divided by U(k,k) in the back substitution).  The pieces and their order
// depend on the pattern only, so a row's arithmetic is the same for every tile width.
// The pieces are list-scheduled onto the G lane groups, kBackItems items per group and step: at
// each step, groups with free item slots take the pieces of the rows whose entries' rows have all
// ended before that step, rows with the longest remaining path first; a group's pieces follow each
// other in its item stream, a piece starting a step later where its END would share a step with the
// group's previous END (one END per group and step), a row's COMBINEs no earlier than the step after
// its partials ended.  Pieces are numbered in the order they end (step, then group); an END stores
// into ring row (number mod (ring_rows - 1)), which the piece's START names; the last ring row holds
// zeros.  An ENTRY of step s reads x_c from its ring row while fewer than ring_rows - 1 pieces ended
// after c's row did before step s, else from y, which the kernel loads at the end of step
// s - kBackAhead: c must have ended by then.  A COMBINE reads its partial from the ring.  Returns
// false (no schedule) when the encoding's limits are exceeded or a read would come too early.
bool build_solve_levels(const gx::LuPattern& lu, int G, int ring_rows, bool forward, std::vector<int>& items,
                        int& steps);

bool build_back_levels(const gx::LuPattern& lu, int G, int ring_rows, std::vector<int>& items, int& steps) {
  return build_solve_levels(lu, G, ring_rows, false, items, steps);
}

// The forward substitution L y = f (L unit lower) in the same item format; END keeps the sum (no
// division; payload 0).
bool build_forward_levels(const gx::LuPattern& lu, int G, int ring_rows, std::vector<int>& items, int& steps) {
  return build_solve_levels(lu, G, ring_rows, true, items, steps);
}

bool build_solve_levels(const gx::LuPattern& lu, int G, int ring_rows, bool forward, std::vector<int>& items,
                        int& steps) {
  const int n = lu.n, P = gx::kBackItems, D = gx::kBackAhead, TW = gx::kLanes / G, S = TW + 1;
  items.clear();
  steps = 0;
  if (n <= 0 || (long long)n * TW >= (1LL << 31) || ring_rows < 2 || (long long)ring_rows * S >= (1LL << 29) ||
      lu.nnz() >= (1 << 29))
    return false;
  const int xrows = ring_rows - 1;                                   // ring rows for x values
  const int zero_b = (int)(0x80000000u | (unsigned)(xrows * S));     // the last ring row holds zeros
  // U by rows: row k -> (c, packed position of U(k, c)), descending c; or L by rows: row k ->
  // (j, position of L(k, j)), ascending j
  auto first = [&](int c) { return forward ? lu.cp[c] + lu.ulen[c] + 1 : lu.cp[c]; };
  auto past = [&](int c) { return forward ? lu.cp[c + 1] : lu.cp[c] + lu.ulen[c]; };
  std::vector<int> rptr(n + 1, 0);
  for (int c = 0; c < n; ++c)
    for (int q = first(c); q < past(c); ++q) rptr[lu.row[q] + 1]++;
  for (int k = 0; k < n; ++k) rptr[k + 1] += rptr[k];
  std::vector<int> rc(rptr[n]), rq(rptr[n]), cur(rptr.begin(), rptr.end() - 1);
  for (int c = 0; c < n; ++c)
    for (int q = first(c); q < past(c); ++q) {
      const int k = lu.row[q];
      rc[cur[k]] = c;
      rq[cur[k]] = q;
      ++cur[k];
    }
  if (!forward)
    for (int k = 0; k < n; ++k) {  // filled in ascending c: reverse each row
      std::reverse(rc.begin() + rptr[k], rc.begin() + rptr[k + 1]);
      std::reverse(rq.begin() + rptr[k], rq.begin() + rptr[k + 1]);
    }
  std::vector<int> lev(n, 0);
  for (int i = 0; i < n; ++i) {
    const int k = forward ? i : n - 1 - i;  // every row after the rows it reads
    int l = 0;
    for (int e = rptr[k]; e < rptr[k + 1]; ++e) l = std::max(l, lev[rc[e]] + 1);
    lev[k] = l;
  }
  // entries of each row by the level of the row they read (ord: indices into rc / rq)
  std::vector<int> ord(rptr[n]);
  for (int k = 0; k < n; ++k) {
    for (int e = rptr[k]; e < rptr[k + 1]; ++e) ord[e] = e;
    std::stable_sort(ord.begin() + rptr[k], ord.begin() + rptr[k + 1], [&](int a, int b) { return lev[rc[a]] < lev[rc[b]]; });
  }
  // pieces: row k's pieces are pfirst[k] .. pfirst[k + 1] - 1, the last one its own; piece v holds
  // ord[pe[v] .. pe[v + 1])
  std::vector<int> pfirst(n + 1, 0), pe, prow;
  for (int k = 0; k < n; ++k) {
    const int L = rptr[k + 1] - rptr[k];
    const int m = L <= gx::kSolvePiece ? 1 : std::min(gx::kSolvePieces, (L + gx::kSolvePiece - 1) / gx::kSolvePiece);
    for (int i = 0, o = rptr[k]; i < m; ++i) {
      pe.push_back(o);
      prow.push_back(k);
      o += L / m + (i < L % m ? 1 : 0);
    }
    pfirst[k + 1] = (int)pe.size();
  }
  const int nv = (int)pe.size();
  pe.push_back(rptr[n]);
  // priority: the longest path (in steps) from a row to the end of the solve
  auto dur = [&](int k) {
    int longest = 0;
    for (int v = pfirst[k]; v < pfirst[k + 1]; ++v) longest = std::max(longest, pe[v + 1] - pe[v]);
    return (longest + 2 + (pfirst[k + 1] - pfirst[k] - 1) + P - 1) / P;
  };
  std::vector<std::vector<int>> users(n);
  for (int k = 0; k < n; ++k)
    for (int e = rptr[k]; e < rptr[k + 1]; ++e) users[rc[e]].push_back(k);
  std::vector<long long> bl(n, 0);
  for (int i = n - 1; i >= 0; --i) {
    const int k = forward ? i : n - 1 - i;
    long long b = 0;
    for (int u : users[k]) b = std::max(b, bl[u]);
    bl[k] = b + dur(k);
  }
  // list schedule.  Cells per group by item position; kind PAD where nothing runs.
  struct Cell {
    int kind = gx::kBackPad, v = 0, e = 0;
  };
  std::vector<std::vector<Cell>> gl(G);
  std::vector<long long> pos(G, 0), last_end(G, -1);
  std::vector<int> end_step(nv, -1), end_group(nv, -1), row_end(n, -1), ndep(n);
  auto put = [&](int g, long long at, Cell c) {
    if ((long long)gl[g].size() <= at) gl[g].resize((size_t)at + 1);
    gl[g][(size_t)at] = c;
  };
  // places piece v on group g no earlier than position lo; COMBINEs (owner pieces) no earlier than
  // step cmin; returns the END's step
  auto place = [&](int v, int g, long long lo, long long cmin) {
    const int k = prow[v], nent = pe[v + 1] - pe[v];
    const bool own = v == pfirst[k + 1] - 1;
    const int ncomb = own ? v - pfirst[k] : 0;
    long long st = std::max(pos[g], lo);
    long long cpos = 0, endp = 0;
    for (;;) {
      cpos = st + 1 + nent;  // first COMBINE (or the END)
      if (ncomb > 0 && cpos < cmin * P) cpos = cmin * P;
      endp = cpos + ncomb;
      if (endp / P == last_end[g]) {
        st = (last_end[g] + 1) * P;
        continue;
      }
      break;
    }
    put(g, st, {gx::kBackStart, v, 0});
    for (int i = 0; i < nent; ++i) put(g, st + 1 + i, {gx::kBackEntry, v, ord[pe[v] + i]});
    for (int i = 0; i < ncomb; ++i) put(g, cpos + i, {gx::kBackCombine, v, pfirst[k] + i});
    put(g, endp, {gx::kBackEnd, v, 0});
    pos[g] = endp + 1;
    last_end[g] = endp / P;
    end_step[v] = (int)(endp / P);
    end_group[v] = g;
    return (int)(endp / P);
  };
  using Ent = std::pair<long long, int>;  // (-priority, row)
  std::priority_queue<Ent, std::vector<Ent>, std::greater<Ent>> ready;
  std::priority_queue<std::pair<int, int>, std::vector<std::pair<int, int>>, std::greater<std::pair<int, int>>> rel;
  for (int k = 0; k < n; ++k) {
    ndep[k] = rptr[k + 1] - rptr[k];
    if (ndep[k] == 0) rel.push({0, k});
  }
  int done = 0;
  long long s = 0;
  while (done < n) {
    while (!rel.empty() && rel.top().first <= s) {
      const int k = rel.top().second;
      rel.pop();
      ready.push({-bl[k], k});
    }
    // the group whose stream is free earliest (ties: the lowest group)
    auto least = [&]() {
      int best = 0;
      for (int g = 1; g < G; ++g)
        if (std::max(pos[g], s * P) < std::max(pos[best], s * P)) best = g;
      return best;
    };
    while (!ready.empty() && pos[least()] < (s + 1) * P) {
      const int k = ready.top().second;
      ready.pop();
      int cmin = 0;
      for (int v = pfirst[k]; v < pfirst[k + 1] - 1; ++v)  // partials, each where a stream is free first
        cmin = std::max(cmin, place(v, least(), s * P, 0) + 1);
      row_end[k] = place(pfirst[k + 1] - 1, least(), s * P, cmin);
      ++done;
      for (int u : users[k])
        if (--ndep[u] == 0) {
          int r = 0;
          for (int e = rptr[u]; e < rptr[u + 1]; ++e) r = std::max(r, row_end[rc[e]] + 1);
          rel.push({r, u});
        }
    }
    ++s;
    if (ready.empty() && !rel.empty()) {  // idle until the next row is released
      bool idle = true;
      for (int g = 0; g < G; ++g) idle = idle && pos[g] <= s * P;
      if (idle) s = std::max<long long>(s, rel.top().first);
    }
    if (s >= (1LL << 26)) return false;
  }
  long long nst = 0;
  for (int g = 0; g < G; ++g) nst = std::max<long long>(nst, ((long long)gl[g].size() + P - 1) / P);
  const long long total = (nst + 2 * D - 1) / (2 * D) * (2 * D);
  if (total * P * G * 2 >= (1LL << 31)) return false;
  // numbering in the order pieces end; ended_before[s] = pieces that ended in steps before s
  std::vector<int> order(nv), fin(nv);
  for (int v = 0; v < nv; ++v) order[v] = v;
  std::sort(order.begin(), order.end(), [&](int a, int b) {
    return end_step[a] != end_step[b] ? end_step[a] < end_step[b] : end_group[a] < end_group[b];
  });
  for (int i = 0; i < nv; ++i) fin[order[i]] = i;
  std::vector<int> ended_before(total + 1, 0);
  for (int v = 0; v < nv; ++v) ended_before[end_step[v] + 1]++;
  for (long long q = 0; q < total; ++q) ended_before[q + 1] += ended_before[q];
  items.assign((size_t)(2 * total * P * G), 0);
  for (long long q = 0; q < total; ++q)
    for (int i = 0; i < P; ++i)
      for (int g = 0; g < G; ++g) {
        const size_t at = (size_t)(q * P + i);
        const Cell c = at < gl[g].size() ? gl[g][at] : Cell();
        const int sq = (int)q;
        int a = gx::kBackPad << 29, b = zero_b;
        const int k = prow[c.v];
        if (c.kind == gx::kBackStart) {
          const bool own = c.v == pfirst[k + 1] - 1;
          a = (gx::kBackStart << 29) | (fin[c.v] % xrows) * S;
          b = own ? k * TW : zero_b;
        } else if (c.kind == gx::kBackEnd) {
          const bool own = c.v == pfirst[k + 1] - 1;
          a = (gx::kBackEnd << 29) | (forward || !own ? 0 : lu.cp[k] + lu.ulen[k]);
        } else if (c.kind == gx::kBackEntry) {
          const int src = rc[c.e], sv = pfirst[src + 1] - 1;  // the source row's own piece
          if (end_step[sv] >= sq) return false;
          a = (gx::kBackEntry << 29) | rq[c.e];
          if (ended_before[sq] <= fin[sv] + xrows - 1) {
            b = (int)(0x80000000u | (unsigned)((fin[sv] % xrows) * S));
          } else {
            if (end_step[sv] > sq - D) return false;  // the kernel loads it at the end of step s - D
            b = src * TW;
          }
        } else if (c.kind == gx::kBackCombine) {
          const int pv = c.e;  // the partial
          if (end_step[pv] >= sq || ended_before[sq] > fin[pv] + xrows - 1) return false;
          a = (int)((unsigned)gx::kBackCombine << 29);
          b = (int)(0x80000000u | (unsigned)((fin[pv] % xrows) * S));
        }
        const size_t x = ((size_t)q * P + i) * G + g;
        items[2 * x] = a;
        items[2 * x + 1] = b;
      }
  steps = (int)total;
  return true;
}

struct DevSched {
  DevMem cp, ulen, row, chunk0, rounds, slots, fptr, fpos, fcol, bptr, bpos, bcol, col_a, col_f, col_b, fslot, opos, ocol;
  DevMem uround, uitems, bslot, bitems, fitems;
  int back_steps = 0, fwd_steps = 0;
  // refactor parts for one lane-group count (they depend on nothing else): [1] holds 32 groups
  // (2-scenario tiles, and quad tiles, whose refactor runs at twice the solves' groups), [0] the
  // others, so a quad-tile batch and the FP64 re-solve of its fallbacks (16 groups) keep both
  struct PartSet {
    DevMem part_a, pent_pos, pent_div, part_x, part_w;
    int groups = -1, n_parts = 0, part_cap = 0;
  } pset[2];
  std::vector<int> h_chunk0, h_rounds, h_fptr, h_bptr;
  bool uploaded = false;
  int window = -1;
  int groups = -1;
  void reset() {
    uploaded = false;
    window = -1;
    groups = -1;
    pset[0].groups = pset[1].groups = -1;
  }
};

// Lanes of an L+U buffer for `lanes` lanes of tiles of `tile`: the tile count rounded up to whole
// blocks of gx::lu_block(tile) tiles (the device layout interleaves a block's tiles).
size_t lu_lanes(size_t lanes, int tile) {
  const size_t tiles = (lanes + tile - 1) / tile, b = (size_t)gx::lu_block(tile);
  return (tiles + b - 1) / b * b * (size_t)tile;
}

// Narrow tiles refactor by parts (lu_tile_parts) when this is set; else column by column.
constexpr bool kUseParts = true;

// Entries of the LU kernel's staging buffer: a refactor part holds up to twice the longest column
// (within the item encoding's limit), at least kPartRegs entries per lane group.  Parts of two
// columns' length: 10,000-bus grid, 8,192 mixed scenarios, 2,685 -> fewer parts, LU 37.0 -> 36.1 ms
// per launch.
int lu_cap(int max_col, int tile) {
  if (!kUseParts || tile == gx::kWideTile) return max_col;
  return std::max(std::max(max_col, std::min(2 * max_col, gx::kPartMaxEntries)), (gx::kLanes / tile) * gx::kPartRegs);
}

// Uploads the pattern once, the update rounds whenever the lane-group count changes, the far
// lists whenever the window changes, and the refactor parts of part_tile (0: tile) lanes per group
// the first time they are asked for.
int upload_sched(const gx::LuPattern& lu, const LuSchedule& s, int window, int tile, DevSched& m,
                 gx::DevTopo& d, const std::vector<int>* kind = nullptr, int part_tile = 0) {
  int rc;
  const int G = gx::kLanes / tile;
  if (part_tile <= 0) part_tile = tile;
  if (!m.uploaded) {
    if ((rc = m.cp.upload(lu.cp)) || (rc = m.ulen.upload(lu.ulen)) || (rc = m.row.upload(lu.row))) return rc;
    m.uploaded = true;
    m.window = -1;
    m.groups = -1;
  }
  const bool regroup = m.groups != G, rewindow = m.window != window;
  if (regroup && tile == gx::kWideTile) {
    std::vector<int> ur, items;
    if ((rc = build_uniform_rounds(lu, s, ur, items)) || (rc = m.uround.upload(ur)) || (rc = m.uitems.upload(items)))
      return rc;
    m.h_chunk0.swap(ur);  // col_a.w = first round of the column
    m.h_rounds.assign(lu.n, 0);
    m.groups = G;
  } else if (regroup) {
    RoundPlan plan;
    if ((rc = build_rounds(lu, s, G, plan)) || (rc = m.chunk0.upload(plan.chunk0)) ||
        (rc = m.rounds.upload(plan.rounds)) || (rc = m.slots.upload(plan.slots)))
      return rc;
    m.h_chunk0.swap(plan.chunk0);
    m.h_rounds.swap(plan.rounds);
    m.groups = G;
  }
  if (rewindow) {
    FarLists f;
    build_far(lu, window, f);
    if ((rc = m.fptr.upload(f.fptr)) || (rc = m.fpos.upload(f.fpos)) || (rc = m.fcol.upload(f.fcol)) ||
        (rc = m.bptr.upload(f.bptr)) || (rc = m.bpos.upload(f.bpos)) || (rc = m.bcol.upload(f.bcol)))
      return rc;
    m.h_fptr.swap(f.fptr);
    m.h_bptr.swap(f.bptr);
    m.window = window;
  }
  if ((regroup || rewindow) && tile == gx::kWideTile) {
    // wide tile: the first kWideFar far entries of the rows entering the forward / backward window
    // after column k in fslot[k][i] / bslot[k][i] ({position, column}, ascending columns), the rest
    // in the ranges col_f = {0, q0, q1, 0} of fwd_pos/fwd_col and col_b = {q0, q1, 0, 0} of
    // bwd_pos/bwd_col
    const int n = lu.n, MF = gx::kWideFar;
    FarLists f;
    build_far(lu, window, f);
    std::vector<int> a(4 * (size_t)n), fw(4 * (size_t)n, 0), bw(4 * (size_t)n, 0);
    std::vector<int> fslot(2 * (size_t)n * MF, -1), bslot(2 * (size_t)n * MF, -1);
    for (int k = 0; k < n; ++k) {
      a[4 * k] = lu.cp[k];
      a[4 * k + 1] = lu.cp[k + 1] - lu.cp[k];
      a[4 * k + 2] = lu.ulen[k];
      a[4 * k + 3] = m.h_chunk0[k];
      const long long rf = (long long)k + window + 1, rb = (long long)k - window - 1;
      if (rf < n) {
        const int q0 = f.fptr[rf], q1 = f.fptr[rf + 1];
        for (int i = 0; i < MF && q0 + i < q1; ++i) {
          fslot[2 * ((size_t)k * MF + i)] = f.fpos[q0 + i];
          fslot[2 * ((size_t)k * MF + i) + 1] = f.fcol[q0 + i];
        }
        fw[4 * k + 1] = std::min(q0 + MF, q1);
        fw[4 * k + 2] = q1;
      }
      if (rb >= 0) {
        const int q0 = f.bptr[rb], q1 = f.bptr[rb + 1];
        for (int i = 0; i < MF && q0 + i < q1; ++i) {
          bslot[2 * ((size_t)k * MF + i)] = f.bpos[q0 + i];
          bslot[2 * ((size_t)k * MF + i) + 1] = f.bcol[q0 + i];
        }
        bw[4 * k] = std::min(q0 + MF, q1);
        bw[4 * k + 1] = q1;
      }
    }
    if ((rc = m.col_a.upload(a)) || (rc = m.col_f.upload(fw)) || (rc = m.col_b.upload(bw)) ||
        (rc = m.fslot.upload(fslot)) || (rc = m.bslot.upload(bslot)))
      return rc;
  } else if (regroup || rewindow) {
    // far L entries of the row entering the forward window after column k: the first G in
    // fslot[k][i] (entry i for lane group i), the others in an overflow list
    const int n = lu.n;
    std::vector<int> fpos, fcol;
    {
      FarLists f;
      build_far(lu, window, f);
      fpos.swap(f.fpos);
      fcol.swap(f.fcol);
    }
    std::vector<int> a(4 * (size_t)n), fw(4 * (size_t)n, 0), bw(4 * (size_t)n, 0);
    std::vector<int> fslot(2 * (size_t)n * G, -1), bslot(2 * (size_t)n * G, -1), opos, ocol, bpos, bcol;
    {
      FarLists f;
      build_far(lu, window, f);
      bpos.swap(f.bpos);
      bcol.swap(f.bcol);
    }
    for (int k = 0; k < n; ++k) {
      a[4 * k] = lu.cp[k];
      a[4 * k + 1] = lu.cp[k + 1] - lu.cp[k];
      a[4 * k + 2] = lu.ulen[k];
      a[4 * k + 3] = m.h_chunk0[k];
      fw[4 * k] = m.h_rounds[k];
      const long long rf = (long long)k + window + 1, rb = (long long)k - window - 1;
      fw[4 * k + 1] = fw[4 * k + 2] = (int)opos.size();
      if (rf < n) {
        for (int q = m.h_fptr[rf]; q < m.h_fptr[rf + 1]; ++q) {
          const int i = q - m.h_fptr[rf];
          if (i < G) {
            fslot[2 * ((size_t)k * G + i)] = fpos[q];
            fslot[2 * ((size_t)k * G + i) + 1] = fcol[q];
          } else {
            opos.push_back(fpos[q]);
            ocol.push_back(fcol[q]);
          }
        }
        fw[4 * k + 2] = (int)opos.size();
        fw[4 * k + 3] = m.h_fptr[rf + 1] > m.h_fptr[rf];
      }
      if (rb >= 0) {
        const int q0 = m.h_bptr[rb], q1 = m.h_bptr[rb + 1];
        for (int i = 0; i < G && q0 + i < q1; ++i) {
          bslot[2 * ((size_t)k * G + i)] = bpos[q0 + i];
          bslot[2 * ((size_t)k * G + i) + 1] = bcol[q0 + i];
        }
        bw[4 * k] = std::min(q0 + G, q1);
        bw[4 * k + 1] = q1;
        bw[4 * k + 2] = q1 > q0;
      }
    }
    if ((rc = m.col_a.upload(a)) || (rc = m.col_f.upload(fw)) || (rc = m.col_b.upload(bw)) ||
        (rc = m.fslot.upload(fslot)) || (rc = m.bslot.upload(bslot)) || (rc = m.opos.upload(opos)) ||
        (rc = m.ocol.upload(ocol)))
      return rc;
    // the back substitution's LDS slots are the solve ring's rows (W + kRingGap + 1)
    std::vector<int> bitems;
    m.back_steps = 0;
    if (build_back_levels(lu, G, window + gx::kRingGap + 1, bitems, m.back_steps)) {
      if ((rc = m.bitems.upload(bitems))) return rc;
    }
    // refactor by parts (below) with the forward substitution by levels
    m.fwd_steps = 0;
    if (kUseParts && m.back_steps > 0) {
      std::vector<int> fitems;
      int fsteps = 0;
      if (build_forward_levels(lu, G, window + gx::kRingGap + 1, fitems, fsteps)) {
        if ((rc = m.fitems.upload(fitems))) return rc;
        m.fwd_steps = fsteps;
      }
    }
  }
  DevSched::PartSet* ps = nullptr;
  if (tile != gx::kWideTile && m.fwd_steps > 0) {
    const int PG = gx::kLanes / part_tile;
    ps = &m.pset[PG == 32 ? 1 : 0];
    if (ps->groups != PG) {
      ps->groups = PG;
      ps->n_parts = 0;
      PartPlan plan;
      if (build_parts(lu, s, PG, lu_cap(s.max_col, part_tile), kind, plan) == GX_OK) {
        // the round items as two streams, x and word, each chunk's items of lane group g one run of
        // NB ints (the kernel loads a chunk with NB / 4 16-byte loads per lane)
        const int NB = gx::part_chunk_rounds(PG);
        const size_t chunks = plan.slots.size() / (2 * (size_t)NB * PG);
        std::vector<int> xs(chunks * PG * NB), ws(chunks * PG * NB);
        for (size_t c = 0; c < chunks; ++c)
          for (int r = 0; r < NB; ++r)
            for (int gg = 0; gg < PG; ++gg) {
              const size_t from = 2 * ((c * NB + r) * PG + gg), to = (c * PG + gg) * NB + r;
              xs[to] = plan.slots[from];
              ws[to] = plan.slots[from + 1];
            }
        if ((rc = ps->part_a.upload(plan.part_a)) || (rc = ps->pent_pos.upload(plan.pent_pos)) ||
            (rc = ps->pent_div.upload(plan.pent_div)) || (rc = ps->part_x.upload(xs)) || (rc = ps->part_w.upload(ws))) {
          ps->groups = -1;
          return rc;
        }
        ps->n_parts = plan.n_parts;
        ps->part_cap = lu_cap(s.max_col, part_tile);
      }
    }
    if (ps->n_parts == 0) ps = nullptr;
  }
  d.n = lu.n;
  d.nnz_lu = lu.nnz();
  d.max_col = s.max_col;
  d.cp = m.cp.as<int>();
  d.ulen = m.ulen.as<int>();
  d.lu_row = m.row.as<int>();
  d.chunk0 = m.chunk0.as<int>();
  d.rounds = m.rounds.as<int>();
  d.slots = m.slots.as<int2>();
  d.groups = G;
  d.col_a = m.col_a.as<int4>();
  d.col_f = m.col_f.as<int4>();
  d.col_b = m.col_b.as<int4>();
  d.fslot = m.fslot.as<int2>();
  d.bslot = m.bslot.as<int2>();
  d.fo_pos = m.opos.as<int>();
  d.fo_col = m.ocol.as<int>();
  d.uround = m.uround.as<int>();
  d.uitems = m.uitems.as<int2>();
  d.fwd_ptr = m.fptr.as<int>();
  d.fwd_pos = m.fpos.as<int>();
  d.fwd_col = m.fcol.as<int>();
  d.bwd_ptr = m.bptr.as<int>();
  d.bwd_pos = m.bpos.as<int>();
  d.bwd_col = m.bcol.as<int>();
  d.window = window;
  d.back_items = tile != gx::kWideTile && m.back_steps > 0 ? m.bitems.as<int2>() : nullptr;
  d.back_steps = d.back_items ? m.back_steps : 0;
  const bool parts = ps != nullptr;
  d.part_a = parts ? ps->part_a.as<int4>() : nullptr;
  d.pent_pos = parts ? ps->pent_pos.as<int>() : nullptr;
  d.pent_div = parts ? ps->pent_div.as<int>() : nullptr;
  d.part_x = parts ? ps->part_x.as<int>() : nullptr;
  d.part_w = parts ? ps->part_w.as<int>() : nullptr;
  d.n_parts = parts ? ps->n_parts : 0;
  d.part_groups = parts ? ps->groups : 0;
  d.stage_zero = parts && kind ? 1 : 0;
  d.fwd_items = parts ? m.fitems.as<int2>() : nullptr;
  d.fwd_steps = parts ? m.fwd_steps : 0;
  if (parts) d.max_col = ps->part_cap;  // (sizes the staging buffers: lu_cap)
  return GX_OK;
}

// Rows of the solve ring for a launch: the requested (or default) power of two, halved while the
// LU kernel's LDS (staged column + ring, `elem` bytes per value) does not fit; -1 if nothing fits.
int pick_ring_rows(int max_col, int tile, size_t elem, int requested) {
  auto bytes = [&](int rows) { return gx::lu_lds_bytes(max_col, rows, tile, elem, true); };
  int rows = requested > 0 ? requested : gx::ring_rows(tile);
  while (rows > 64 && bytes(rows) > (size_t)gx::kMaxLds) rows /= 2;
  return bytes(rows) <= (size_t)gx::kMaxLds ? rows : -1;
}

bool valid_tile(int tile) { return tile == 2 || tile == 4 || tile == 8 || tile == 16 || tile == gx::kWideTile; }
bool valid_ring(int rows) { return rows == 0 || (rows >= 64 && rows <= 4096 && (rows & (rows - 1)) == 0); }

// Scenarios per tile when the caller leaves it open: 8-scenario pair tiles once they make 1,024
// wavefronts (one per SIMD; 512 in mixed precision); else 2-scenario tiles while their wavefronts fit one per SIMD, else
// the narrowest of 4 and 8 whose wavefronts stay within 2,048 per launch, else 16.  Measured (LU + fill per Newton step): with 1,024 scenarios on the 10,000-bus grid
// 2-scenario tiles beat 4-scenario tiles (27.9 vs 36.4 ms LU at FP64); on the 2,000-bus grid with
// 4,096 scenarios the 2,048 wavefronts of 2-scenario tiles lose (16.1 vs 10.3 ms); on the
// 10,000-bus grid with 8,192 scenarios 4-scenario tiles (two wavefronts per SIMD, run one after the
// other at the kernel's register count) beat 8-scenario tiles at FP64 (96 + 11 vs 137 + 8 ms) and
// at FP32 (54 + 15 vs 76 + 5 ms).
// With at least 1,024 tiles of 8 scenarios, pair tiles of 8 (two scenarios per lane on the
// 4-scenario schedule, one wavefront per SIMD), mixed precision: 10,000-bus grid, 8,192 scenarios, LU
// 37.0 against 39.4 ms per launch, fill 4.1 against 6.2 ms, batch 895 against 993 ms.  Their
// refactor runs as quad tiles (four scenarios per lane on the 2-scenario part schedule): 29.4
// against 31.4 ms per launch.  FP64 with 8,192 scenarios: pair tiles 60.5 ms per LU launch against
// 62.0 for 4-scenario tiles and 82.9 for 8-scenario tiles of one scenario per lane.  Mixed precision
// takes quad tiles from 512 of them, although half the SIMDs then idle: 2,000-bus grid, 4,096
// scenarios, 150.0 against 162.5 ms per batch (FP64 pair tiles there: 159.8 against 154.1 ms).
int auto_tile(int n_scen, bool fp32) {
  if ((n_scen + 7) / 8 >= (fp32 ? 512 : 1024)) return 8;
  if ((n_scen + 1) / 2 <= 1024) return 2;
  for (int tw : {4, 8}) if ((n_scen + tw - 1) / tw <= 2048) return tw;
  return 16;
}

// Position of (pivot row r, pivot column k) inside the packed column k, or -1.
int lu_position(const gx::LuPattern& lu, int r, int k) {
  const int c0 = lu.cp[k], ul = lu.ulen[k];
  if (r == k) return c0 + ul;
  if (r < k) {
    auto b = lu.row.begin() + c0, e = lu.row.begin() + c0 + ul;
    auto it = std::lower_bound(b, e, r);
    return (it != e && *it == r) ? (int)(it - lu.row.begin()) : -1;
  }
  auto b = lu.row.begin() + c0 + ul + 1, e = lu.row.begin() + lu.cp[k + 1];
  auto it = std::lower_bound(b, e, r);
  return (it != e && *it == r) ? (int)(it - lu.row.begin()) : -1;
}

// Triplets -> CSC with rows sorted per column; slot[e] = CSC position of triplet e
// (build_csc_structure, src/klu_native/mod.rs:88-114).
void triplets_to_csc(int n, const std::vector<int>& rows, const std::vector<int>& cols, gx::CscPattern& a,
                     std::vector<int>& slot) {
  const size_t nt = rows.size();
  std::vector<int> ord(nt);
  for (size_t i = 0; i < nt; ++i) ord[i] = (int)i;
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
    return cols[x] != cols[y] ? cols[x] < cols[y] : rows[x] < rows[y];
  });
  a.n = n;
  a.ap.assign(n + 1, 0);
  a.ai.clear();
  slot.assign(nt, -1);
  for (size_t t = 0; t < nt; ++t) {
    int e = ord[t];
    if (t > 0 && rows[e] == rows[ord[t - 1]] && cols[e] == cols[ord[t - 1]]) {
      slot[e] = (int)a.ai.size() - 1;
      continue;
    }
    a.ai.push_back(rows[e]);
    a.ap[cols[e] + 1]++;
    slot[e] = (int)a.ai.size() - 1;
  }
  for (int j = 0; j < n; ++j) a.ap[j + 1] += a.ap[j];
}

// The four Y-bus contributions of one branch (build_ybus, src/network.rs:225-247; transformers
// by branch_calc_param (1,1), src/network.rs:339-360): y[0..3] = ff, ft, tf, tt.  A line from a
// bus to itself is a pure shunt (ff only).
struct BranchStamp {
  int f = 0, t = 0;
  bool transformer = false;
  cx y[4];
};

int branch_stamps(int n_bus, int n_line, const int* lf, const int* lt, const double* r, const double* x,
                  const double* bsh, const double* gsh, int n_tr, const int* tf, const int* tt, const double* ys,
                  const double* ysh, const double* tap, std::vector<BranchStamp>& out) {
  out.clear();
  out.reserve((size_t)n_line + n_tr);
  for (int e = 0; e < n_line; ++e) {
    if (lf[e] < 0 || lf[e] >= n_bus || lt[e] < 0 || lt[e] >= n_bus) return fail(GX_ERR_ARG, "line bus out of range");
    BranchStamp b;
    b.f = lf[e];
    b.t = lt[e];
    if (lf[e] == lt[e]) {
      b.y[0] = {gsh ? gsh[e] : 0.0, bsh[e]};
      b.y[1] = b.y[2] = b.y[3] = {0.0, 0.0};
    } else {
      cx yl = cdiv({1.0, 0.0}, {r[e], x[e]});
      cx half{(gsh ? gsh[e] : 0.0) / 2.0, bsh[e] / 2.0};
      b.y[0] = cadd(yl, half);
      b.y[3] = cadd(yl, half);
      b.y[1] = cneg(yl);
      b.y[2] = cneg(yl);
    }
    out.push_back(b);
  }
  for (int e = 0; e < n_tr; ++e) {
    if (tf[e] < 0 || tf[e] >= n_bus || tt[e] < 0 || tt[e] >= n_bus) return fail(GX_ERR_ARG, "transformer bus out of range");
    cx s{ys[2 * e], ys[2 * e + 1]}, sh{ysh[2 * e], ysh[2 * e + 1]}, a{tap[2 * e], tap[2 * e + 1]};
    double k = std::hypot(a.re, a.im);
    cx yd = cadd(s, {sh.re * 0.5, sh.im * 0.5});
    BranchStamp b;
    b.f = tf[e];
    b.t = tt[e];
    b.transformer = true;
    b.y[0] = cdiv(yd, {k * k, 0.0});
    b.y[1] = cdiv(cneg(s), cconj(a));
    b.y[2] = cdiv(cneg(s), a);
    b.y[3] = yd;
    out.push_back(b);
  }
  return GX_OK;
}

// Bridges of the branch multigraph (self-loops never are): removing one splits an island.
std::vector<char> find_bridges(int n_bus, const std::vector<BranchStamp>& br) {
  std::vector<std::vector<std::pair<int, int>>> adj(n_bus);
  for (int e = 0; e < (int)br.size(); ++e)
    if (br[e].f != br[e].t) {
      adj[br[e].f].push_back({br[e].t, e});
      adj[br[e].t].push_back({br[e].f, e});
    }
  std::vector<char> bridge(br.size(), 0);
  std::vector<int> disc(n_bus, -1), low(n_bus, 0);
  int timer = 0;
  // iterative DFS: (vertex, edge used to enter, next adjacency index)
  std::vector<std::tuple<int, int, size_t>> st;
  for (int r0 = 0; r0 < n_bus; ++r0) {
    if (disc[r0] >= 0) continue;
    disc[r0] = low[r0] = timer++;
    st.emplace_back(r0, -1, 0);
    while (!st.empty()) {
      auto& [v, pe, idx] = st.back();
      if (idx < adj[v].size()) {
        auto [u, e] = adj[v][idx++];
        if (e == pe) continue;
        if (disc[u] < 0) {
          disc[u] = low[u] = timer++;
          st.emplace_back(u, e, 0);
        } else {
          low[v] = std::min(low[v], disc[u]);
        }
      } else {
        const int vv = v, ee = pe;
        st.pop_back();
        if (!st.empty()) {
          const int parent = std::get<0>(st.back());
          low[parent] = std::min(low[parent], low[vv]);
          if (low[vv] > disc[parent]) bridge[ee] = 1;
        }
      }
    }
  }
  return bridge;
}

// Device buffers of one batch: the buffers of gx_batch_solve (one per topology, reused across
// calls) or of an engine request.
struct Workspace {
  size_t lanes = 0;
  int iter_cap = 0;
  int fp32 = -1;
  DevMem vm, va, vr, vi, ps, qs, pc, qc, f, dx, y, lu, status, iters, maxmis, hist, running, dvr, dvi, stage,
      fallback, vm0, va0, outage, iter_dev, lg_lu, lg_y, lg_status, lg_iters, swept, sing0;
  int* h_running = nullptr;  // pinned host mirror of the running counters
  int h_running_len = 0;
  Workspace() = default;
  Workspace(const Workspace&) = delete;
  Workspace& operator=(const Workspace&) = delete;
  ~Workspace() {
    if (h_running) (void)hipHostFree(h_running);
  }
  size_t bytes() const {
    const DevMem* all[] = {&vm, &va, &vr, &vi, &ps, &qs, &pc, &qc, &f, &dx, &y, &lu, &status, &iters, &maxmis,
                           &hist, &running, &dvr, &dvi, &stage, &fallback, &vm0, &va0, &outage, &iter_dev,
                           &lg_lu, &lg_y, &lg_status, &lg_iters, &swept, &sing0};
    size_t b = 0;
    for (const DevMem* m : all) b += m->bytes;
    return b;
  }
};

}  // namespace

struct gx_topo {
  int n_bus = 0;
  std::vector<int> type;        // effective bus types (unreferenced islands marked Slack)
  std::vector<int> given_type;  // bus types as passed in
  // islands (connected components of the Y-bus graph, numbered by lowest bus)
  std::vector<int> comp;        // bus -> component
  std::vector<int> verdict;     // component -> GX_ISLAND_*
  std::vector<int> dev_type;    // device bus codes: 0/1/2, 3 = member of an unreferenced island
  DevMem d_comp, w_island;
  std::vector<int> yrow, ycol;
  std::vector<double> yg, yb;
  std::vector<int> non_slack, pq, ns_pos, pq_pos;
  int n_ang = 0, n = 0;
  // Jacobian recipe in reference emission order
  std::vector<int> jrow, jcol, jkind, jbi, jbk;
  std::vector<double> jg, jb;
  gx::CscPattern csc;
  std::vector<int> slot;
  gx::Ordering ord;
  gx::LuPattern lu;
  LuSchedule sched;
  bool factored = false;
  // host images of the device schedule (uploaded lazily by upload_topo)
  std::vector<int> h_kind, h_bi, h_bk, h_rp, h_rq, h_xa, h_xm;
  std::vector<double> h_g, h_b;
  bool uploaded = false;
  // device topology
  DevMem d_yrow, d_ycol, d_yg, d_yb, d_type, d_rpos_p, d_rpos_q, d_xpos_a, d_xpos_m;
  DevMem d_kind, d_bi, d_bk, d_g, d_b;
  DevSched ds;
  gx::DevTopo dt;
  // launch shape: scenarios per tile and solve ring rows (0 = automatic)
  int tile_req = 0, ring_req = 0;
  int refactor_groups = 0;  // lane groups of the FP32 8-scenario refactor: 0 automatic, 16 or 32
  // workspace of gx_batch_solve
  Workspace ws;
  int last_fallback = 0;
  // linear initial guess system (built on first use)
  bool lg_built = false;
  gx::LuPattern lg_lu;
  LuSchedule lg_sched;
  std::vector<int> lg_src, lg_dbus, lg_row_bus, lg_row_piv, lg_row_x, lg_rhs_ptr, lg_rhs_e, lg_rhs_j;
  DevMem d_lg_src, d_lg_dbus, d_lg_row_bus, d_lg_row_piv, d_lg_row_x, d_lg_rhs_ptr, d_lg_rhs_e, d_lg_rhs_j;
  DevSched lg_ds;
  gx::DevLg dlg;
  gx::DevTopo dlgt;
  bool ybus_uploaded = false;
  // resident warm start: |V| then theta per bus, and a device flag set once it holds a solution
  DevMem d_warm, d_warm_ok;
  // ZIP terms (gx_topo_set_zip): 6 per bus, empty when no bus has any
  std::vector<double> zip;
  DevMem d_zip;
  bool zip_uploaded = true;
  // branches for per-scenario outages (gx_topo_set_branches)
  int n_br = 0;
  std::vector<int> br_pos;     // 4 per branch: Y CSR positions of ff, ft, tf, tt (-1 if none)
  std::vector<double> br_y;    // 8 per branch
  std::vector<char> br_bridge;
  std::vector<int> h_yp;       // per L+U position: the Y CSR position its recipe reads (-1 for fill)
  // recipes by bus pair (DevTopo::fb_*) and the fill positions
  std::vector<int> h_fb_ptr, h_fb_k, h_fb_pos, h_fb_yp, h_fz, h_ylu;
  std::vector<double> h_fb_g, h_fb_b;
  DevMem d_fb_ptr, d_fb_k, d_fb_pos, d_fb_g, d_fb_b, d_fb_yp, d_fz, d_ylu;
  DevMem d_br_pos, d_br_y, d_yp;
  bool br_uploaded = false;
  // the last gx_batch_solve (for island reports)
  gx::DevWork last;
  bool has_last = false;
  // opt-in per-launch timing: (kernel class, start, stop)
  bool profiling = false;
  std::vector<std::tuple<int, hipEvent_t, hipEvent_t>> prof;
  // events of the pending timings come from a pool that each finished solve refills; the timings
  // are summed per kernel class until gx_topo_profile reads them
  std::vector<hipEvent_t> ev_pool;
  double prof_ms[GX_K_COUNT] = {};
  int prof_n[GX_K_COUNT] = {};
  ~gx_topo() {
    for (auto& e : prof) {
      (void)hipEventDestroy(std::get<1>(e));
      (void)hipEventDestroy(std::get<2>(e));
    }
    for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
  }
};

namespace {
// Sums the pending timings (their events must have completed) and returns the events to the pool.
int fold_profile(gx_topo& t) {
  for (auto& e : t.prof) {
    float v = 0.0f;
    GX_HIP(hipEventSynchronize(std::get<2>(e)));
    GX_HIP(hipEventElapsedTime(&v, std::get<1>(e), std::get<2>(e)));
    t.prof_ms[std::get<0>(e)] += v;
    t.prof_n[std::get<0>(e)] += 1;
    t.ev_pool.push_back(std::get<1>(e));
    t.ev_pool.push_back(std::get<2>(e));
  }
  t.prof.clear();
  return GX_OK;
}
}  // namespace

namespace {

// Position of Y(i, j) in the CSR Y-bus (rows sorted by column), or -1.
int ybus_position(const gx_topo& t, int i, int j) {
  auto b = t.ycol.begin() + t.yrow[i], e = t.ycol.begin() + t.yrow[i + 1];
  auto it = std::lower_bound(b, e, j);
  return (it != e && *it == j) ? (int)(it - t.ycol.begin()) : -1;
}

// Host FP64 injections for one scenario (used only to fix the pivot order).
void host_injections(const gx_topo& t, const double* vm, const double* va, std::vector<double>& p,
                     std::vector<double>& q) {
  const int n = t.n_bus;
  std::vector<cx> v(n);
  for (int i = 0; i < n; ++i) v[i] = {vm[i] * std::cos(va[i]), vm[i] * std::sin(va[i])};
  p.assign(n, 0.0);
  q.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    cx acc{0.0, 0.0};
    for (int e = t.yrow[i]; e < t.yrow[i + 1]; ++e) acc = cadd(acc, cmul({t.yg[e], t.yb[e]}, v[t.ycol[e]]));
    cx s = cmul(v[i], cconj(acc));
    p[i] = s.re;
    q[i] = s.im;
  }
}

// Jacobian values in reference emission order at one operating point (src/jacobian.rs:210-250).
void host_jacobian(const gx_topo& t, const double* vm, const double* va, const std::vector<double>& p,
                   const std::vector<double>& q, std::vector<double>& vals) {
  vals.resize(t.jkind.size());
  for (size_t e = 0; e < t.jkind.size(); ++e) {
    const int i = t.jbi[e], k = t.jbk[e];
    const double g = t.jg[e], b = t.jb[e], vi = vm[i];
    double v = 0.0;
    switch (t.jkind[e]) {
      case gx::kHii: v = -q[i] - vi * vi * b; break;
      case gx::kNii: v = p[i] / vi + vi * g; break;
      case gx::kMii: v = p[i] - vi * vi * g; break;
      case gx::kLii: v = q[i] / vi - vi * b; break;
      default: {
        const double vk = vm[k], ang = va[i] - va[k], sn = std::sin(ang), cs = std::cos(ang);
        switch (t.jkind[e]) {
          case gx::kHik: v = vi * vk * (g * sn - b * cs); break;
          case gx::kNik: v = vi * (g * cs + b * sn); break;
          case gx::kMik: v = -vi * vk * (g * cs + b * sn); break;
          default: v = vi * (g * sn - b * cs); break;
        }
      }
    }
    vals[e] = v;
  }
}

// Derives the device schedule (host images) once the pivot order is known.
int finalize_factor(gx_topo& t) {
  int rc = build_schedule(t.lu, t.sched);
  if (rc) return rc;
  const gx::LuPattern& lu = t.lu;
  const int nnz = lu.nnz();
  std::vector<int> qinv(t.n);
  for (int k = 0; k < t.n; ++k) qinv[lu.q[k]] = k;
  std::vector<int> kind(nnz, gx::kFill), bi(nnz, 0), bk(nnz, 0), yp(nnz, -1);
  std::vector<double> g(nnz, 0.0), b(nnz, 0.0);
  for (size_t e = 0; e < t.jkind.size(); ++e) {
    int pos = lu_position(lu, lu.pinv[t.jrow[e]], qinv[t.jcol[e]]);
    if (pos < 0) return fail(GX_ERR_ARG, "Jacobian entry missing from the L+U pattern");
    if (kind[pos] != gx::kFill) return fail(GX_ERR_UNSUPPORTED, "duplicate Jacobian entries (unmerged Y-bus)");
    kind[pos] = t.jkind[e];
    yp[pos] = ybus_position(t, t.jbi[e], t.jbk[e]);
    bi[pos] = t.jbi[e];
    bk[pos] = t.jbk[e];
    g[pos] = t.jg[e];
    b[pos] = t.jb[e];
  }
  std::vector<int> rp(t.n_bus, -1), rq(t.n_bus, -1), xa(t.n_bus, -1), xm(t.n_bus, -1);
  for (int i = 0; i < t.n_bus; ++i) {
    if (t.ns_pos[i] >= 0) {
      rp[i] = lu.pinv[t.ns_pos[i]];
      xa[i] = qinv[t.ns_pos[i]];
    }
    if (t.pq_pos[i] >= 0) {
      rq[i] = lu.pinv[t.n_ang + t.pq_pos[i]];
      xm[i] = qinv[t.n_ang + t.pq_pos[i]];
    }
  }
  // the fill kernel's recipes: one entry per (row bus, column bus) pair with its up to four positions
  {
    std::map<std::pair<int, int>, int> pair_of;
    std::vector<std::array<int, 4>> pp;
    std::vector<int> pk, pi, pyp;
    std::vector<double> pg, pb;
    std::vector<int> fz;
    for (int pos = 0; pos < nnz; ++pos) {
      if (kind[pos] == gx::kFill) {
        fz.push_back(pos);
        continue;
      }
      const auto key = std::make_pair(bi[pos], bk[pos]);
      auto it = pair_of.find(key);
      int e;
      if (it == pair_of.end()) {
        e = (int)pp.size();
        pair_of.emplace(key, e);
        pp.push_back({-1, -1, -1, -1});
        pi.push_back(bi[pos]);
        pk.push_back(bk[pos]);
        pg.push_back(g[pos]);
        pb.push_back(b[pos]);
        pyp.push_back(yp[pos]);
      } else {
        e = it->second;
      }
      const int kd = kind[pos];
      const int slot = (kd == gx::kHii || kd == gx::kHik) ? 0 : (kd == gx::kNii || kd == gx::kNik) ? 1
                       : (kd == gx::kMii || kd == gx::kMik) ? 2 : 3;
      pp[e][slot] = pos;
    }
    // by row bus (map order: row bus, then column bus)
    std::vector<int> fptr(t.n_bus + 1, 0), fk, fpos, fyp;
    std::vector<double> fg, fb;
    for (auto& kv : pair_of) fptr[kv.first.first + 1]++;
    for (int i = 0; i < t.n_bus; ++i) fptr[i + 1] += fptr[i];
    for (auto& kv : pair_of) {
      const int e = kv.second;
      fk.push_back(pk[e]);
      for (int c = 0; c < 4; ++c) fpos.push_back(pp[e][c]);
      fg.push_back(pg[e]);
      fb.push_back(pb[e]);
      fyp.push_back(pyp[e]);
    }
    t.h_fb_ptr.swap(fptr);
    t.h_fb_k.swap(fk);
    t.h_fb_pos.swap(fpos);
    t.h_fb_g.swap(fg);
    t.h_fb_b.swap(fb);
    t.h_fb_yp.swap(fyp);
    // the same positions by Y-bus position, for the mismatch kernel that fills while it walks the
    // Y rows (empty, so the separate fill runs, if a pair has no Y position)
    std::vector<int> ylu(4 * t.ycol.size(), -1);
    bool whole = true;
    for (size_t e = 0; e < t.h_fb_yp.size(); ++e) {
      const int p = t.h_fb_yp[e];
      if (p < 0 || (size_t)p >= t.ycol.size()) {
        whole = false;
        break;
      }
      for (int c = 0; c < 4; ++c) ylu[4 * (size_t)p + c] = t.h_fb_pos[4 * e + c];
    }
    if (!whole) ylu.clear();
    t.h_ylu.swap(ylu);
    t.h_fz.swap(fz);
  }
  t.h_kind.swap(kind);
  t.h_yp.swap(yp);
  t.h_bi.swap(bi);
  t.h_bk.swap(bk);
  t.h_g.swap(g);
  t.h_b.swap(b);
  t.h_rp.swap(rp);
  t.h_rq.swap(rq);
  t.h_xa.swap(xa);
  t.h_xm.swap(xm);
  t.factored = true;
  t.uploaded = false;
  t.ds.reset();
  return GX_OK;
}

int upload_ybus(gx_topo& t);

// Uploads the topology (Y-bus, recipes, LU pattern and schedule) once per factorization, and the
// far lists of the solves for the given window.
int upload_topo(gx_topo& t, int window, int tile, int part_tile = 0) {
  int rc;
  if ((rc = upload_sched(t.lu, t.sched, window, tile, t.ds, t.dt, &t.h_kind, part_tile))) return rc;
  if (t.uploaded) return GX_OK;
  if ((rc = upload_ybus(t)) || (rc = t.d_rpos_p.upload(t.h_rp)) ||
      (rc = t.d_rpos_q.upload(t.h_rq)) || (rc = t.d_xpos_a.upload(t.h_xa)) || (rc = t.d_xpos_m.upload(t.h_xm)) ||
      (rc = t.d_kind.upload(t.h_kind)) || (rc = t.d_bi.upload(t.h_bi)) || (rc = t.d_bk.upload(t.h_bk)) ||
      (rc = t.d_g.upload(t.h_g)) || (rc = t.d_b.upload(t.h_b)) || (rc = t.d_yp.upload(t.h_yp)) ||
      (rc = t.d_fb_ptr.upload(t.h_fb_ptr)) || (rc = t.d_fb_k.upload(t.h_fb_k)) ||
      (rc = t.d_fb_pos.upload(t.h_fb_pos)) || (rc = t.d_fb_g.upload(t.h_fb_g)) ||
      (rc = t.d_fb_b.upload(t.h_fb_b)) || (rc = t.d_fb_yp.upload(t.h_fb_yp)) || (rc = t.d_fz.upload(t.h_fz)) ||
      (rc = t.d_ylu.upload(t.h_ylu)))
    return rc;
  gx::DevTopo& d = t.dt;
  d.fb_ptr = t.d_fb_ptr.as<int>();
  d.fb_k = t.d_fb_k.as<int>();
  d.fb_pos = t.d_fb_pos.as<int4>();
  d.fb_g = t.d_fb_g.as<double>();
  d.fb_b = t.d_fb_b.as<double>();
  d.fb_yp = t.d_fb_yp.as<int>();
  d.fz_pos = t.d_fz.as<int>();
  d.y_lupos = t.h_ylu.empty() ? nullptr : t.d_ylu.as<int4>();
  d.n_fz = (int)t.h_fz.size();
  d.lu_yp = t.d_yp.as<int>();
  d.rpos_p = t.d_rpos_p.as<int>();
  d.rpos_q = t.d_rpos_q.as<int>();
  d.xpos_a = t.d_xpos_a.as<int>();
  d.xpos_m = t.d_xpos_m.as<int>();
  d.lu_kind = t.d_kind.as<int>();
  d.lu_bi = t.d_bi.as<int>();
  d.lu_bk = t.d_bk.as<int>();
  d.lu_g = t.d_g.as<double>();
  d.lu_b = t.d_b.as<double>();
  t.uploaded = true;
  return GX_OK;
}

// Per-batch buffers that do not depend on the factorization (the L+U buffer is reserved after it).
int ensure_workspace(const gx_topo& t, Workspace& ws, size_t lanes, int max_iter, bool fp32) {
  if (lanes <= ws.lanes && max_iter <= ws.iter_cap && (int)fp32 == ws.fp32) return GX_OK;
  const size_t nb = lanes * t.n_bus * sizeof(double);
  const size_t nu = lanes * t.n * sizeof(double);
  int rc;
  if ((rc = ws.vm.reserve(nb)) || (rc = ws.va.reserve(nb)) || (rc = ws.vr.reserve(nb)) || (rc = ws.vi.reserve(nb)) ||
      (rc = ws.ps.reserve(nb)) || (rc = ws.qs.reserve(nb)) || (rc = ws.pc.reserve(nb)) || (rc = ws.qc.reserve(nb)) ||
      (rc = ws.f.reserve(nu)) || (rc = ws.dx.reserve(nu)) || (rc = ws.y.reserve(nu)) ||
      (rc = ws.status.reserve(lanes * sizeof(int))) || (rc = ws.iters.reserve(lanes * sizeof(int))) ||
      (rc = ws.maxmis.reserve(lanes * sizeof(unsigned long long))) ||
      (rc = ws.hist.reserve(lanes * (size_t)std::max(max_iter, 1) * sizeof(double))) ||
      (rc = ws.running.reserve((size_t)std::max(max_iter, 1) * sizeof(int))) ||
      (rc = ws.fallback.reserve(sizeof(int))))
    return rc;
  if (ws.h_running_len < std::max(max_iter, 1)) {
    if (ws.h_running) (void)hipHostFree(ws.h_running);
    ws.h_running = nullptr;
    ws.h_running_len = 0;
    GX_HIP(hipHostMalloc((void**)&ws.h_running, sizeof(int) * std::max(max_iter, 1), hipHostMallocDefault));
    ws.h_running_len = std::max(max_iter, 1);
  }
  if (fp32 && ((rc = ws.dvr.reserve(nb)) || (rc = ws.dvi.reserve(nb)) || (rc = ws.swept.reserve(lanes * sizeof(int)))))
    return rc;
  ws.lanes = lanes;
  ws.iter_cap = max_iter;
  ws.fp32 = (int)fp32;
  return GX_OK;
}

int upload_branches(gx_topo& t) {
  if (t.br_uploaded || t.n_br == 0) return GX_OK;
  int rc;
  if ((rc = t.d_br_pos.upload(t.br_pos)) || (rc = t.d_br_y.upload(t.br_y))) return rc;
  t.dt.br_pos = t.d_br_pos.as<int4>();
  t.dt.br_y = t.d_br_y.as<double>();
  t.br_uploaded = true;
  return GX_OK;
}

int upload_ybus(gx_topo& t) {
  int rc;
  if (!t.zip_uploaded) {
    if (t.zip.empty()) {
      t.dt.zip = nullptr;
    } else {
      if ((rc = t.d_zip.upload(t.zip))) return rc;
      t.dt.zip = t.d_zip.as<double>();
    }
    t.zip_uploaded = true;
  }
  if (t.ybus_uploaded) return GX_OK;
  if ((rc = t.d_yrow.upload(t.yrow)) || (rc = t.d_ycol.upload(t.ycol)) || (rc = t.d_yg.upload(t.yg)) ||
      (rc = t.d_yb.upload(t.yb)) || (rc = t.d_type.upload(t.dev_type)) || (rc = t.d_comp.upload(t.comp)))
    return rc;
  gx::DevTopo& d = t.dt;
  d.n_bus = t.n_bus;
  d.n = t.n;
  d.y_rowptr = t.d_yrow.as<int>();
  d.y_col = t.d_ycol.as<int>();
  d.y_g = t.d_yg.as<double>();
  d.y_b = t.d_yb.as<double>();
  d.bus_type = t.d_type.as<int>();
  t.ybus_uploaded = true;
  return GX_OK;
}

// The linear initial guess system (src/network.rs:627-678) for the bus types as given: the
// reference runs the guess before islands are marked (src/batch.rs:268, src/solver.rs:452-454).
int build_lg(gx_topo& t, Workspace& ws, size_t lanes, int tile) {
  int rc;
  if (!t.lg_built) {
    std::vector<int> pq, pos(t.n_bus, -1);
    for (int i = 0; i < t.n_bus; ++i)
      if (t.given_type[i] == 2) {
        pos[i] = (int)pq.size();
        pq.push_back(i);
      }
    const int m = (int)pq.size();
    std::vector<int> tr, tc, tsrc, tdbus;
    t.lg_rhs_ptr.assign(m + 1, 0);
    t.lg_rhs_e.clear();
    t.lg_rhs_j.clear();
    for (int r = 0; r < m; ++r) {
      const int i = pq[r];
      bool diag = false;
      for (int e = t.yrow[i]; e < t.yrow[i + 1]; ++e) {
        const int j = t.ycol[e];
        if (j == i) diag = true;
        if (pos[j] >= 0) {
          tr.push_back(r);
          tc.push_back(pos[j]);
          tsrc.push_back(e);
          tdbus.push_back(j == i ? i : -1);
        } else {
          t.lg_rhs_e.push_back(e);
          t.lg_rhs_j.push_back(j);
        }
      }
      if (!diag) {
        tr.push_back(r);
        tc.push_back(r);
        tsrc.push_back(-1);
        tdbus.push_back(i);
      }
      t.lg_rhs_ptr[r + 1] = (int)t.lg_rhs_e.size();
    }
    t.dlg.m = m;
    if (m > 0) {
      gx::CscPattern a;
      std::vector<int> slot;
      triplets_to_csc(m, tr, tc, a, slot);
      gx::Ordering ord;
      if (!gx::analyze(a, ord) || !gx::pattern_from_order(a, ord.q, ord.p, t.lg_lu))
        return fail(GX_ERR_SINGULAR, "linear initial guess system is structurally singular");
      if ((rc = build_schedule(t.lg_lu, t.lg_sched))) return rc;
      std::vector<int> qinv(m);
      for (int k = 0; k < m; ++k) qinv[t.lg_lu.q[k]] = k;
      const int nlu = t.lg_lu.nnz();
      t.lg_src.assign(nlu, -1);
      t.lg_dbus.assign(nlu, -1);
      for (size_t e = 0; e < tr.size(); ++e) {
        int p = lu_position(t.lg_lu, t.lg_lu.pinv[tr[e]], qinv[tc[e]]);
        if (p < 0) return fail(GX_ERR_ARG, "linear guess entry missing from its L+U pattern");
        t.lg_src[p] = tsrc[e];
        t.lg_dbus[p] = tdbus[e];
      }
      t.lg_row_bus = pq;
      t.lg_row_piv.assign(m, 0);
      t.lg_row_x.assign(m, 0);
      for (int r = 0; r < m; ++r) {
        t.lg_row_piv[r] = t.lg_lu.pinv[r];
        t.lg_row_x[r] = qinv[r];
      }
      if ((rc = t.d_lg_src.upload(t.lg_src)) || (rc = t.d_lg_dbus.upload(t.lg_dbus)) ||
          (rc = t.d_lg_row_bus.upload(t.lg_row_bus)) || (rc = t.d_lg_row_piv.upload(t.lg_row_piv)) ||
          (rc = t.d_lg_row_x.upload(t.lg_row_x)) || (rc = t.d_lg_rhs_ptr.upload(t.lg_rhs_ptr)) ||
          (rc = t.d_lg_rhs_e.upload(t.lg_rhs_e)) || (rc = t.d_lg_rhs_j.upload(t.lg_rhs_j)))
        return rc;
      gx::DevLg& l = t.dlg;
      l.nnz_lu = nlu;
      l.src = t.d_lg_src.as<int>();
      l.dbus = t.d_lg_dbus.as<int>();
      l.row_bus = t.d_lg_row_bus.as<int>();
      l.row_piv = t.d_lg_row_piv.as<int>();
      l.row_x = t.d_lg_row_x.as<int>();
      l.rhs_ptr = t.d_lg_rhs_ptr.as<int>();
      l.rhs_e = t.d_lg_rhs_e.as<int>();
      l.rhs_j = t.d_lg_rhs_j.as<int>();
    }
    t.lg_built = true;
  }
  if (t.dlg.m > 0) {
    const int rows = pick_ring_rows(lu_cap(t.lg_sched.max_col, tile), tile, 16, 0);
    if (rows < 0) return fail(GX_ERR_UNSUPPORTED, "linear guess L+U column exceeds the LDS staging limit");
    if ((rc = upload_sched(t.lg_lu, t.lg_sched, gx::ring_window_of(rows, tile), tile, t.lg_ds, t.dlgt))) return rc;
    if ((rc = ws.lg_lu.reserve(lu_lanes(lanes, tile) * (size_t)t.dlg.nnz_lu * 16)) || (rc = ws.lg_y.reserve(lanes * t.dlg.m * 16)) ||
        (rc = ws.lg_status.reserve(lanes * sizeof(int))) || (rc = ws.lg_iters.reserve(lanes * sizeof(int))))
      return rc;
  }
  return GX_OK;
}

int factor_at(gx_topo& t, const double* vm, const double* va, const double* ps, const double* qs) {
  (void)ps;
  (void)qs;
  std::vector<double> p, q, vals;
  host_injections(t, vm, va, p, q);
  host_jacobian(t, vm, va, p, q, vals);
  std::vector<double> ax(t.csc.ai.size(), 0.0);
  for (size_t e = 0; e < vals.size(); ++e) ax[t.slot[e]] += vals[e];
  if (!gx::factor_pivot(t.csc, ax.data(), t.ord, 1e-3, t.lu))
    return fail(GX_ERR_SINGULAR, "first factorization found a zero or missing pivot");
  return finalize_factor(t);
}

// One branch out of service per scenario (-1: none), validated; bridges would split an island.
int read_outages(const gx_topo& t, int n_scen, const int* outage, bool dev, std::vector<int>& h, bool& any) {
  any = false;
  if (!outage) return GX_OK;
  h.resize(n_scen);
  if (dev) GX_HIP(hipMemcpy(h.data(), outage, sizeof(int) * n_scen, hipMemcpyDeviceToHost));
  else std::memcpy(h.data(), outage, sizeof(int) * n_scen);
  for (int s = 0; s < n_scen; ++s) {
    const int e = h[s];
    if (e < -1 || e >= t.n_br)
      return fail(t.n_br == 0 ? GX_ERR_STATE : GX_ERR_ARG,
                  "scenario " + std::to_string(s) + ": outage branch out of range (gx_topo_set_branches first)");
    if (e >= 0 && t.br_bridge[e])
      return fail(GX_ERR_ARG, "scenario " + std::to_string(s) + ": branch " + std::to_string(e) +
                                  " is a bridge; its outage splits an island");
    any = any || e >= 0;
  }
  return GX_OK;
}

struct BatchIn {
  int n_scen = 0;
  const double *p = nullptr, *q = nullptr, *vm0 = nullptr, *va0 = nullptr;
  double tol = 0.0;
  int max_iter = 0;
  bool fp32 = false;
  int flags = 0;
  bool dev = false;
  hipStream_t st = nullptr;
  const std::vector<int>* outage = nullptr;  // validated host copy, or null when none is out
};

// Results of a batch for which no scenario's first Jacobian could be factored (nothing runs on the
// device): every scenario converged at its starting point or is Singular after its first mismatch,
// with the starting point as its voltages.
struct HostResult {
  bool used = false;
  std::vector<double> vm, va, mismatch;  // [n_scen][n_bus], [n_scen]
};

// Stages a batch on the device: inputs in tiles, optional linear initial guess, the topology's first
// factorization (if not yet factored), the topology upload for this tile width, the init kernel and
// the mixed-precision starting-point copy.  Fills w.  When no scenario's first Jacobian can be
// factored, fills *hres instead (GX_ERR_SINGULAR when hres is null).
// The resident warm-start buffers of a topology (allocated and marked empty on first use).
int ensure_warm(gx_topo& t) {
  if (t.d_warm.p) return GX_OK;
  int rc;
  if ((rc = t.d_warm.reserve(2 * (size_t)t.n_bus * sizeof(double))) || (rc = t.d_warm_ok.reserve(sizeof(int))))
    return rc;
  GX_HIP(hipMemset(t.d_warm_ok.p, 0, sizeof(int)));
  return GX_OK;
}

int stage_batch(gx_topo& t, Workspace& ws, const BatchIn& in, int tile, gx::DevWork& w, HostResult* hres,
                bool pairs = false, bool quads = false) {
  const int n_scen = in.n_scen, nb = t.n_bus, max_iter = in.max_iter;
  const size_t plane = (size_t)n_scen * nb * sizeof(double);
  hipStream_t st = in.st;
  const int n_tiles = (n_scen + tile - 1) / tile;
  const size_t lanes = (size_t)n_tiles * tile;
  int rc;
  if ((rc = upload_ybus(t)) || (rc = ensure_workspace(t, ws, lanes, max_iter, in.fp32))) return rc;
  w = gx::DevWork();
  w.tile = tile;
  w.n_scen = n_scen;
  w.n_tiles = n_tiles;
  w.max_iter = max_iter;
  w.tol = in.tol;
  w.vm = ws.vm.as<double>();
  w.va = ws.va.as<double>();
  w.vr = ws.vr.as<double>();
  w.vi = ws.vi.as<double>();
  w.ps = ws.ps.as<double>();
  w.qs = ws.qs.as<double>();
  w.pc = ws.pc.as<double>();
  w.qc = ws.qc.as<double>();
  w.f = ws.f.as<double>();
  w.dx = ws.dx.as<double>();
  w.y = ws.y.p;
  w.status = ws.status.as<int>();
  w.iters = ws.iters.as<int>();
  w.maxmis = ws.maxmis.as<unsigned long long>();
  w.history = ws.hist.as<double>();
  w.running = ws.running.as<int>();
  w.fallback = ws.fallback.as<int>();
  if (in.outage) {
    if ((rc = upload_branches(t)) || (rc = ws.outage.upload(*in.outage))) return rc;
    w.outage = ws.outage.as<int>();
  }
  if ((in.flags & GX_INIT_WARM) && (in.flags & GX_INIT_LINEAR))
    return fail(GX_ERR_ARG, "GX_INIT_WARM and GX_INIT_LINEAR exclude each other");
  // inputs -> interleaved tiles
  const double* src[4] = {in.p, in.q, in.vm0, in.va0};
  double* dst[4] = {w.ps, w.qs, w.vm, w.va};
  if ((rc = ws.stage.reserve(plane))) return rc;
  for (int a = 0; a < 4; ++a) {
    const double* s = src[a];
    if (!in.dev) {
      GX_HIP(hipMemcpyAsync(ws.stage.p, src[a], plane, hipMemcpyHostToDevice, st));
      s = ws.stage.as<double>();
    }
    gx::launch_to_tiles(s, dst[a], n_scen, n_tiles, nb, tile, st);
  }
  if (in.flags & GX_INIT_WARM) {
    if ((rc = ensure_warm(t))) return rc;
    gx::launch_warm_start(t.dt, w, t.d_warm.as<double>(), t.d_warm_ok.as<int>(), st);
  }
  if (in.flags & GX_INIT_LINEAR) {
    // BatchSolver runs linear_initial_guess on every scenario first (src/batch.rs:268)
    if ((rc = build_lg(t, ws, lanes, tile))) return rc;
    gx::launch_linear_guess(t.dt, t.dlg, t.dlgt, w, ws.lg_lu.p, ws.lg_y.p, ws.lg_status.as<int>(),
                            ws.lg_iters.as<int>(), st);
    GX_HIP(hipGetLastError());
  }
  if (!t.factored) {
    // PersistentSolver's first factorization uses the first Jacobian it is asked to solve: the
    // first scenario whose starting point is not already converged (src/solver.rs:699-713).  When
    // that Jacobian has no valid pivot order the scenario is Singular and the cache stays empty, so
    // the next such scenario factors its own (src/solver.rs:711-716).
    gx::launch_from_tiles(w.vm, ws.stage.as<double>(), n_scen, nb, tile, st);
    std::vector<double> hvm((size_t)n_scen * nb);
    GX_HIP(hipMemcpyAsync(hvm.data(), ws.stage.p, plane, hipMemcpyDeviceToHost, st));
    GX_HIP(hipStreamSynchronize(st));
    gx::launch_from_tiles(w.va, ws.stage.as<double>(), n_scen, nb, tile, st);
    std::vector<double> hva((size_t)n_scen * nb);
    GX_HIP(hipMemcpyAsync(hva.data(), ws.stage.p, plane, hipMemcpyDeviceToHost, st));
    GX_HIP(hipStreamSynchronize(st));
    std::vector<double> hp(nb), hq(nb), pc, qc, mis(n_scen, 0.0);
    std::vector<int> failed;  // scenarios whose first Jacobian could not be factored
    int s = 0;
    for (; s < n_scen && !t.factored; ++s) {
      const size_t off = (size_t)s * nb;
      hipMemcpyKind k = in.dev ? hipMemcpyDeviceToHost : hipMemcpyHostToHost;
      GX_HIP(hipMemcpy(hp.data(), in.p + off, nb * sizeof(double), k));
      GX_HIP(hipMemcpy(hq.data(), in.q + off, nb * sizeof(double), k));
      for (int i = 0; i < nb; ++i)
        if (t.dev_type[i] == gx::kDeadBus) hvm[off + i] = hva[off + i] = hp[i] = hq[i] = 0.0;
      host_injections(t, hvm.data() + off, hva.data() + off, pc, qc);
      if (!t.zip.empty())
        for (int i = 0; i < nb; ++i) {
          const double v = hvm[off + i], *z = &t.zip[6 * (size_t)i];
          hp[i] = hp[i] + z[0] + z[2] * v + z[4] * v * v;
          hq[i] = hq[i] + z[1] + z[3] * v + z[5] * v * v;
        }
      double mx = 0.0;
      for (int i = 0; i < nb; ++i) {
        if (t.ns_pos[i] >= 0) mx = std::fmax(mx, std::fabs(hp[i] - pc[i]));
        if (t.pq_pos[i] >= 0) mx = std::fmax(mx, std::fabs(hq[i] - qc[i]));
      }
      mis[s] = mx;
      const bool converged = mx < in.tol;
      // (a batch that starts converged still fixes a pivot order, at its last scenario, so that the
      // device has one to run with)
      if (converged && s < n_scen - 1) continue;
      rc = factor_at(t, hvm.data() + off, hva.data() + off, hp.data(), hq.data());
      if (rc != GX_OK && rc != GX_ERR_SINGULAR) return rc;
      if (rc == GX_ERR_SINGULAR && !converged) failed.push_back(s);
    }
    if (!t.factored) {
      // no valid pivot order in any scenario: the results are the starting points
      if (!hres) return fail(GX_ERR_SINGULAR, "no scenario's first Jacobian has a valid pivot order");
      hres->used = true;  // (every scenario was examined, its dead buses zeroed)
      hres->vm.swap(hvm);
      hres->va.swap(hva);
      hres->mismatch.swap(mis);
      return GX_OK;
    }
    if (!failed.empty()) {
      std::vector<int> flag(lanes, 0);
      for (int q : failed) flag[q] = 1;
      if ((rc = ws.sing0.upload(flag))) return rc;
      w.sing0 = ws.sing0.as<int>();
    }
  }
  // the solve ring is sized for FP64 values: the mixed mode's FP64 re-solve uses the same lists.
  // Pair tiles run the schedules of tiles half as wide (two scenarios per lane).
  // Quad tiles refactor on lane groups half as wide again (four scenarios per lane), when their
  // part schedule exists; else they run as pair tiles.
  int lt = pairs ? tile / 2 : tile;
  int rows = pick_ring_rows(lu_cap(t.sched.max_col, lt), lt, (pairs ? 2 : 1) * sizeof(double), t.ring_req);
  if (pairs && rows < 0) {
    // two values per lane do not fit the LDS with this topology's longest column: one per lane
    pairs = quads = false;
    lt = tile;
    rows = pick_ring_rows(lu_cap(t.sched.max_col, lt), lt, sizeof(double), t.ring_req);
  }
  if (rows < 0)
    return fail(GX_ERR_UNSUPPORTED, "longest L+U column exceeds the LDS staging limit of the device refactor");
  int window = gx::ring_window_of(rows, lt);
  if ((rc = upload_topo(t, window, lt, quads ? lt / 2 : lt))) return rc;
  if (quads && t.dt.part_groups != 2 * (gx::kLanes / lt) && (rc = upload_topo(t, window, lt, lt))) return rc;
  if (pairs && !t.dt.part_a) {
    // pair tiles need the refactor by parts; without a part schedule (a column past the parts'
    // 10-bit offsets, or solves beyond the level schedule's encoding) the tile runs as tile lanes
    lt = tile;
    rows = pick_ring_rows(lu_cap(t.sched.max_col, lt), lt, sizeof(double), t.ring_req);
    if (rows < 0)
      return fail(GX_ERR_UNSUPPORTED, "longest L+U column exceeds the LDS staging limit of the device refactor");
    window = gx::ring_window_of(rows, lt);
    if ((rc = upload_topo(t, window, lt))) return rc;
  }
  // FP32 values for the mixed mode (its FP64 re-solve of fallbacks grows the buffer first)
  if ((rc = ws.lu.reserve(lu_lanes(lanes, tile) * (size_t)t.dt.nnz_lu * (in.fp32 ? 4 : 8)))) return rc;
  w.lu = ws.lu.p;
  if (in.fp32) w.swept = ws.swept.as<int>();
  gx::launch_init(t.dt, w, st);
  GX_HIP(hipGetLastError());
  if (in.fp32) {
    // the FP64 re-solve's starting point, taken after k_init pinned unreferenced islands to V = 0
    const size_t nbytes = lanes * nb * sizeof(double);
    if ((rc = ws.vm0.reserve(nbytes)) || (rc = ws.va0.reserve(nbytes))) return rc;
    w.vm0 = ws.vm0.as<double>();
    w.va0 = ws.va0.as<double>();
    GX_HIP(hipMemcpyAsync(w.vm0, w.vm, nbytes, hipMemcpyDeviceToDevice, st));
    GX_HIP(hipMemcpyAsync(w.va0, w.va, nbytes, hipMemcpyDeviceToDevice, st));
  }
  return GX_OK;
}

// Copies a finished batch's results out (host or device pointers) and waits for them.
int collect_batch(const gx_topo& t, Workspace& ws, const gx::DevWork& w, bool dev, double* vm, double* va,
                  int* status, int* iters, double* history, hipStream_t st) {
  const int n_scen = w.n_scen, nb = t.n_bus;
  const size_t plane = (size_t)n_scen * nb * sizeof(double);
  double* outs[2] = {vm, va};
  const double* tiles[2] = {w.vm, w.va};
  for (int a = 0; a < 2; ++a) {
    if (dev) {
      gx::launch_from_tiles(tiles[a], outs[a], n_scen, nb, w.tile, st);
    } else {
      gx::launch_from_tiles(tiles[a], ws.stage.as<double>(), n_scen, nb, w.tile, st);
      GX_HIP(hipMemcpyAsync(outs[a], ws.stage.p, plane, hipMemcpyDeviceToHost, st));
    }
  }
  hipMemcpyKind k = dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  GX_HIP(hipMemcpyAsync(status, w.status, sizeof(int) * n_scen, k, st));
  GX_HIP(hipMemcpyAsync(iters, w.iters, sizeof(int) * n_scen, k, st));
  if (history && w.max_iter > 0)
    GX_HIP(hipMemcpyAsync(history, w.history, sizeof(double) * n_scen * (size_t)w.max_iter, k, st));
  GX_HIP(hipStreamSynchronize(st));
  return GX_OK;
}

// Writes a HostResult as a batch's outputs (host or device pointers as in.dev).
int host_results(gx_topo& t, const BatchIn& in, const HostResult& h, double* vm, double* va, int* status, int* iters,
                 double* history) {
  const int n_scen = in.n_scen, nb = t.n_bus;
  std::vector<int> st(n_scen), it(n_scen, 1);
  std::vector<double> hist((size_t)n_scen * std::max(in.max_iter, 0), 0.0);
  for (int s = 0; s < n_scen; ++s) {
    st[s] = h.mismatch[s] < in.tol ? GX_CONVERGED : GX_SINGULAR;
    if (in.max_iter > 0) hist[(size_t)s * in.max_iter] = h.mismatch[s];
  }
  if (in.max_iter == 0)
    for (int s = 0; s < n_scen; ++s) {
      st[s] = GX_MAX_ITER;
      it[s] = 0;
    }
  const hipMemcpyKind k = in.dev ? hipMemcpyHostToDevice : hipMemcpyHostToHost;
  const size_t plane = (size_t)n_scen * nb * sizeof(double);
  GX_HIP(hipMemcpy(vm, h.vm.data(), plane, k));
  GX_HIP(hipMemcpy(va, h.va.data(), plane, k));
  GX_HIP(hipMemcpy(status, st.data(), sizeof(int) * n_scen, k));
  GX_HIP(hipMemcpy(iters, it.data(), sizeof(int) * n_scen, k));
  if (history && !hist.empty()) GX_HIP(hipMemcpy(history, hist.data(), sizeof(double) * hist.size(), k));
  t.has_last = false;
  return GX_OK;
}

}  // namespace

extern "C" {

const char* gx_version(void) { return "gridox_hip 0.1.0 (gfx950)"; }

int gx_topo_profile(gx_topo* t, int enable, double* ms, int* launches) {
  if (!t) return fail(GX_ERR_ARG, "null");
  int rc;
  if ((rc = fold_profile(*t))) return rc;
  for (int c = 0; c < GX_K_COUNT; ++c) {
    if (ms) ms[c] = t->prof_ms[c];
    if (launches) launches[c] = t->prof_n[c];
    t->prof_ms[c] = 0.0;
    t->prof_n[c] = 0;
  }
  t->profiling = enable != 0;
  return GX_OK;
}
const char* gx_last_error(void) { return g_err.c_str(); }

int gx_topo_set_launch(gx_topo* t, int tile, int ring_rows) {
  if (!t) return fail(GX_ERR_ARG, "null");
  if (!(tile == 0 || valid_tile(tile))) return fail(GX_ERR_ARG, "tile must be 0, 4, 8 or 16");
  if (!valid_ring(ring_rows)) return fail(GX_ERR_ARG, "ring_rows must be 0 or a power of two in [64, 4096]");
  t->tile_req = tile;
  t->ring_req = ring_rows;
  return GX_OK;
}

int gx_topo_set_refactor_groups(gx_topo* t, int groups) {
  if (!t) return fail(GX_ERR_ARG, "null");
  if (groups != 0 && groups != 16 && groups != 32) return fail(GX_ERR_ARG, "refactor groups must be 0, 16 or 32");
  t->refactor_groups = groups;
  return GX_OK;
}

int gx_ybus_build(int n_bus, int n_line, const int* lf, const int* lt, const double* r, const double* x,
                  const double* bsh, const double* gsh, int n_tr, const int* tf, const int* tt, const double* ys,
                  const double* ysh, const double* tap, int cap, int* rowptr, int* col, double* y_re, double* y_im) {
  if (n_bus <= 0 || n_line < 0 || n_tr < 0 || !rowptr) return fail(GX_ERR_ARG, "bad sizes");
  std::vector<BranchStamp> br;
  int rc = branch_stamps(n_bus, n_line, lf, lt, r, x, bsh, gsh, n_tr, tf, tt, ys, ysh, tap, br);
  if (rc) return rc;
  // insertion-ordered accumulation per (row, col)
  std::map<std::pair<int, int>, cx> acc;
  auto add = [&](int i, int j, cx v) {
    auto it = acc.find({i, j});
    if (it == acc.end()) acc.emplace(std::make_pair(i, j), cadd({0.0, 0.0}, v));
    else it->second = cadd(it->second, v);
  };
  for (const BranchStamp& b : br) {
    add(b.f, b.f, b.y[0]);
    if (b.f == b.t) continue;
    if (b.transformer) {  // stamp order of stamp_transformers: ff, ft, tf, tt
      add(b.f, b.t, b.y[1]);
      add(b.t, b.f, b.y[2]);
      add(b.t, b.t, b.y[3]);
    } else {  // line order: ff, tt, ft, tf
      add(b.t, b.t, b.y[3]);
      add(b.f, b.t, b.y[1]);
      add(b.t, b.f, b.y[2]);
    }
  }
  int nnz = (int)acc.size();
  if (nnz > cap || !col || !y_re || !y_im) return -nnz;
  for (int i = 0; i <= n_bus; ++i) rowptr[i] = 0;
  int t = 0;
  for (auto& kv : acc) {
    rowptr[kv.first.first + 1]++;
    col[t] = kv.first.second;
    y_re[t] = kv.second.re;
    y_im[t] = kv.second.im;
    ++t;
  }
  for (int i = 0; i < n_bus; ++i) rowptr[i + 1] += rowptr[i];
  return nnz;
}

gx_topo* gx_topo_create(int n_bus, const int* bus_type, const int* rowptr, const int* col, const double* y_re,
                        const double* y_im, int* err) {
  auto set = [&](int c) {
    if (err) *err = c;
  };
  if (n_bus <= 0 || !bus_type || !rowptr || !col || !y_re || !y_im) {
    set(fail(GX_ERR_ARG, "null or empty topology"));
    return nullptr;
  }
  std::unique_ptr<gx_topo> t(new gx_topo());
  t->n_bus = n_bus;
  t->type.assign(bus_type, bus_type + n_bus);
  const int nnz = rowptr[n_bus];
  t->yrow.assign(rowptr, rowptr + n_bus + 1);
  t->ycol.assign(col, col + nnz);
  t->yg.assign(y_re, y_re + nnz);
  t->yb.assign(y_im, y_im + nnz);
  for (int i = 0; i < n_bus; ++i) {
    if (t->type[i] < 0 || t->type[i] > 2) {
      set(fail(GX_ERR_ARG, "bus type must be 0 (Slack), 1 (PV) or 2 (PQ)"));
      return nullptr;
    }
    for (int e = rowptr[i]; e < rowptr[i + 1]; ++e)
      if (col[e] < 0 || col[e] >= n_bus || (e > rowptr[i] && col[e] <= col[e - 1])) {
        set(fail(GX_ERR_ARG, "Y-bus rows must be sorted, unique and in range"));
        return nullptr;
      }
  }
  // islands: connected_components / classify / mark_unreferenced_islands (src/network.rs:121-223)
  t->given_type = t->type;
  t->comp.assign(n_bus, -1);
  {
    std::vector<int> stack;
    int nc = 0;
    for (int s0 = 0; s0 < n_bus; ++s0) {
      if (t->comp[s0] >= 0) continue;
      t->comp[s0] = nc;
      stack.assign(1, s0);
      while (!stack.empty()) {
        int i = stack.back();
        stack.pop_back();
        for (int e = rowptr[i]; e < rowptr[i + 1]; ++e) {
          int j = col[e];
          if (j != i && t->comp[j] < 0) {
            t->comp[j] = nc;
            stack.push_back(j);
          }
        }
      }
      ++nc;
    }
    std::vector<int> slacks(nc, 0);
    for (int i = 0; i < n_bus; ++i)
      if (t->type[i] == 0) slacks[t->comp[i]]++;
    t->verdict.assign(nc, GX_ISLAND_SOLVABLE);
    for (int c = 0; c < nc; ++c)
      t->verdict[c] = slacks[c] == 0 ? GX_ISLAND_NO_REFERENCE
                                     : (slacks[c] == 1 ? GX_ISLAND_SOLVABLE : GX_ISLAND_AMBIGUOUS);
    t->dev_type = t->type;
    for (int i = 0; i < n_bus; ++i)
      if (t->verdict[t->comp[i]] == GX_ISLAND_NO_REFERENCE) {
        t->type[i] = 0;
        t->dev_type[i] = 3;
      }
  }
  // unknowns: angle of every non-slack bus, then magnitude of every PQ bus (bus order)
  t->ns_pos.assign(n_bus, -1);
  t->pq_pos.assign(n_bus, -1);
  for (int i = 0; i < n_bus; ++i)
    if (t->type[i] != 0) {
      t->ns_pos[i] = (int)t->non_slack.size();
      t->non_slack.push_back(i);
    }
  for (int i = 0; i < n_bus; ++i)
    if (t->type[i] == 2) {
      t->pq_pos[i] = (int)t->pq.size();
      t->pq.push_back(i);
    }
  t->n_ang = (int)t->non_slack.size();
  t->n = t->n_ang + (int)t->pq.size();
  if (t->n == 0) {
    set(fail(GX_ERR_ARG, "no unknowns (every bus is Slack)"));
    return nullptr;
  }
  // JacobianPattern::analyze emission order (src/jacobian.rs:128-168)
  auto push = [&](int r, int c, int kind, int i, int k, int e) {
    t->jrow.push_back(r);
    t->jcol.push_back(c);
    t->jkind.push_back(kind);
    t->jbi.push_back(i);
    t->jbk.push_back(k);
    t->jg.push_back(y_re[e]);
    t->jb.push_back(y_im[e]);
  };
  for (int ri = 0; ri < t->n_ang; ++ri) {
    int i = t->non_slack[ri];
    for (int e = rowptr[i]; e < rowptr[i + 1]; ++e) {
      int k = col[e];
      if (k == i) {
        push(ri, ri, gx::kHii, i, i, e);
        if (t->pq_pos[i] >= 0) push(ri, t->n_ang + t->pq_pos[i], gx::kNii, i, i, e);
        continue;
      }
      if (t->ns_pos[k] >= 0) push(ri, t->ns_pos[k], gx::kHik, i, k, e);
      if (t->pq_pos[k] >= 0) push(ri, t->n_ang + t->pq_pos[k], gx::kNik, i, k, e);
    }
  }
  for (int ri = 0; ri < (int)t->pq.size(); ++ri) {
    int i = t->pq[ri];
    for (int e = rowptr[i]; e < rowptr[i + 1]; ++e) {
      int k = col[e];
      if (k == i) {
        push(t->n_ang + ri, t->ns_pos[i], gx::kMii, i, i, e);
        push(t->n_ang + ri, t->n_ang + ri, gx::kLii, i, i, e);
        continue;
      }
      if (t->ns_pos[k] >= 0) push(t->n_ang + ri, t->ns_pos[k], gx::kMik, i, k, e);
      if (t->pq_pos[k] >= 0) push(t->n_ang + ri, t->n_ang + t->pq_pos[k], gx::kLik, i, k, e);
    }
  }
  triplets_to_csc(t->n, t->jrow, t->jcol, t->csc, t->slot);
  if (!gx::analyze(t->csc, t->ord)) {
    set(fail(GX_ERR_SINGULAR, "Jacobian pattern is structurally singular"));
    return nullptr;
  }
  set(GX_OK);
  return t.release();
}

void gx_topo_destroy(gx_topo* t) { delete t; }

int gx_batch_fallbacks(const gx_topo* t) {
  if (!t) return fail(GX_ERR_ARG, "null");
  return t->last_fallback;
}

int gx_topo_islands(const gx_topo* t, int* comp, int* verdict) {
  if (!t) return fail(GX_ERR_ARG, "null");
  if (comp) std::memcpy(comp, t->comp.data(), sizeof(int) * t->comp.size());
  if (verdict) std::memcpy(verdict, t->verdict.data(), sizeof(int) * t->verdict.size());
  return (int)t->verdict.size();
}

int gx_batch_island_status(gx_topo* t, int* out) {
  if (!t || !out) return fail(GX_ERR_ARG, "null");
  if (!t->has_last) return fail(GX_ERR_STATE, "no batch solved yet");
  const gx::DevWork& w = t->last;
  const int nc = (int)t->verdict.size();
  const size_t cells = (size_t)w.n_scen * nc;
  int rc;
  if ((rc = t->w_island.reserve(cells * sizeof(unsigned long long)))) return rc;
  GX_HIP(hipMemset(t->w_island.p, 0, cells * sizeof(unsigned long long)));
  gx::launch_island_max(t->dt, w, t->d_comp.as<int>(), nc, t->w_island.as<unsigned long long>(), nullptr);
  GX_HIP(hipGetLastError());
  std::vector<unsigned long long> mx(cells);
  std::vector<int> st(w.n_scen);
  GX_HIP(hipMemcpy(mx.data(), t->w_island.p, cells * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  GX_HIP(hipMemcpy(st.data(), w.status, sizeof(int) * w.n_scen, hipMemcpyDeviceToHost));
  for (int s = 0; s < w.n_scen; ++s)
    for (int c = 0; c < nc; ++c) {
      int v;
      if (t->verdict[c] == GX_ISLAND_NO_REFERENCE) v = GX_ISLAND_STATUS_NO_REFERENCE;
      else if (t->verdict[c] == GX_ISLAND_AMBIGUOUS) v = GX_ISLAND_STATUS_AMBIGUOUS;
      else {
        double m;
        std::memcpy(&m, &mx[(size_t)s * nc + c], sizeof(double));
        v = m < w.tol ? GX_CONVERGED : (st[s] == GX_SINGULAR ? GX_SINGULAR : GX_MAX_ITER);
      }
      out[(size_t)s * nc + c] = v;
    }
  return GX_OK;
}

int gx_topo_sizes(const gx_topo* t, long long* s) {
  if (!t || !s) return fail(GX_ERR_ARG, "null");
  s[0] = t->n_bus;
  s[1] = (long long)t->ycol.size();
  s[2] = t->n;
  s[3] = (long long)t->jrow.size();
  s[4] = t->factored ? t->lu.nnz() : 0;
  s[5] = (long long)t->ord.r.size() - 1;
  s[6] = t->factored ? 1 : 0;
  s[7] = t->factored ? t->sched.n_upd : 0;
  return GX_OK;
}

int gx_topo_jacobian_pattern(const gx_topo* t, int* rows, int* cols) {
  if (!t || !rows || !cols) return fail(GX_ERR_ARG, "null");
  std::memcpy(rows, t->jrow.data(), sizeof(int) * t->jrow.size());
  std::memcpy(cols, t->jcol.data(), sizeof(int) * t->jcol.size());
  return GX_OK;
}

int gx_topo_factor_at(gx_topo* t, const double* vm, const double* va, const double* p_spec, const double* q_spec) {
  if (!t || !vm || !va) return fail(GX_ERR_ARG, "null");
  return factor_at(*t, vm, va, p_spec, q_spec);
}

int gx_topo_set_pivot_order(gx_topo* t, const int* q, const int* pnum) {
  if (!t || !q || !pnum) return fail(GX_ERR_ARG, "null");
  std::vector<int> qq(q, q + t->n), pp(pnum, pnum + t->n);
  if (!gx::pattern_from_order(t->csc, qq, pp, t->lu)) return fail(GX_ERR_ARG, "invalid or structurally singular pivot order");
  return finalize_factor(*t);
}

int gx_topo_ordering(const gx_topo* t, int* p, int* q, int* r, int* pnum) {
  if (!t) return fail(GX_ERR_ARG, "null");
  if (p) std::memcpy(p, t->ord.p.data(), sizeof(int) * t->ord.p.size());
  if (q) std::memcpy(q, t->ord.q.data(), sizeof(int) * t->ord.q.size());
  if (r) std::memcpy(r, t->ord.r.data(), sizeof(int) * t->ord.r.size());
  if (pnum) {
    if (!t->factored) return fail(GX_ERR_STATE, "not factored yet");
    std::memcpy(pnum, t->lu.pnum.data(), sizeof(int) * t->lu.pnum.size());
  }
  return GX_OK;
}

int gx_topo_lu_pattern(const gx_topo* t, int* cp, int* ulen, int* row) {
  if (!t) return fail(GX_ERR_ARG, "null");
  if (!t->factored) return fail(GX_ERR_STATE, "not factored yet");
  if (cp) std::memcpy(cp, t->lu.cp.data(), sizeof(int) * t->lu.cp.size());
  if (ulen) std::memcpy(ulen, t->lu.ulen.data(), sizeof(int) * t->lu.ulen.size());
  if (row) std::memcpy(row, t->lu.row.data(), sizeof(int) * t->lu.row.size());
  return GX_OK;
}

int gx_topo_back_levels(const gx_topo* t, int tile, int ring_rows, int* items, long long cap, int* steps) {
  if (!t || (tile != 2 && tile != 4 && tile != 8 && tile != 16) || ring_rows < 2) return fail(GX_ERR_ARG, "bad arguments");
  if (!t->factored) return fail(GX_ERR_STATE, "not factored yet");
  std::vector<int> it;
  int st = 0;
  if (!build_back_levels(t->lu, gx::kLanes / tile, ring_rows, it, st))
    return fail(GX_ERR_UNSUPPORTED, "system beyond the level schedule's encoding");
  if (steps) *steps = st;
  if ((long long)it.size() > cap || !items) return -(int)std::min<size_t>(it.size(), 0x7fffffff);
  std::memcpy(items, it.data(), sizeof(int) * it.size());
  return (int)it.size();
}

int gx_topo_forward_levels(const gx_topo* t, int tile, int ring_rows, int* items, long long cap, int* steps) {
  if (!t || (tile != 2 && tile != 4 && tile != 8 && tile != 16) || ring_rows < 2) return fail(GX_ERR_ARG, "bad arguments");
  if (!t->factored) return fail(GX_ERR_STATE, "not factored yet");
  std::vector<int> it;
  int st = 0;
  if (!build_forward_levels(t->lu, gx::kLanes / tile, ring_rows, it, st))
    return fail(GX_ERR_UNSUPPORTED, "system beyond the level schedule's encoding");
  if (steps) *steps = st;
  if ((long long)it.size() > cap || !items) return -(int)std::min<size_t>(it.size(), 0x7fffffff);
  std::memcpy(items, it.data(), sizeof(int) * it.size());
  return (int)it.size();
}

int gx_topo_refactor_parts(const gx_topo* t, int tile, long long* sizes, int* part_a, int* pent_pos, int* pent_div,
                           int* slots) {
  if (!t || (tile != 2 && tile != 4 && tile != 8 && tile != 16) || !sizes) return fail(GX_ERR_ARG, "bad arguments");
  if (!t->factored) return fail(GX_ERR_STATE, "not factored yet");
  PartPlan plan;
  const int pcap = lu_cap(t->sched.max_col, tile);
  int rc = build_parts(t->lu, t->sched, gx::kLanes / tile, pcap, &t->h_kind, plan);
  if (rc) return rc;
  sizes[0] = plan.n_parts;
  sizes[1] = (long long)plan.pent_pos.size();
  sizes[2] = (long long)plan.slots.size();
  sizes[3] = pcap;
  if (part_a && pent_pos && pent_div && slots) {
    std::memcpy(part_a, plan.part_a.data(), sizeof(int) * plan.part_a.size());
    std::memcpy(pent_pos, plan.pent_pos.data(), sizeof(int) * plan.pent_pos.size());
    std::memcpy(pent_div, plan.pent_div.data(), sizeof(int) * plan.pent_div.size());
    std::memcpy(slots, plan.slots.data(), sizeof(int) * plan.slots.size());
  }
  return GX_OK;
}

int gx_topo_set_zip(gx_topo* t, const double* zip) {
  if (!t) return fail(GX_ERR_ARG, "null topology");
  if (zip) {
    for (size_t k = 0; k < 6 * (size_t)t->n_bus; ++k)
      if (!std::isfinite(zip[k])) return fail(GX_ERR_ARG, "non-finite ZIP coefficient");
    t->zip.assign(zip, zip + 6 * (size_t)t->n_bus);
  } else {
    t->zip.clear();
  }
  t->zip_uploaded = false;
  return GX_OK;
}

int gx_topo_set_branches(gx_topo* t, int n_line, const int* lf, const int* lt, const double* r, const double* x,
                         const double* bsh, const double* gsh, int n_tr, const int* tf, const int* tt,
                         const double* ys, const double* ysh, const double* tap) {
  if (!t || n_line < 0 || n_tr < 0) return fail(GX_ERR_ARG, "bad branch list");
  std::vector<BranchStamp> br;
  int rc = branch_stamps(t->n_bus, n_line, lf, lt, r, x, bsh, gsh, n_tr, tf, tt, ys, ysh, tap, br);
  if (rc) return rc;
  std::vector<int> pos(4 * br.size(), -1);
  std::vector<double> y(8 * br.size(), 0.0);
  for (size_t e = 0; e < br.size(); ++e) {
    const BranchStamp& b = br[e];
    const int rc4[4][2] = {{b.f, b.f}, {b.f, b.t}, {b.t, b.f}, {b.t, b.t}};
    for (int q = 0; q < (b.f == b.t ? 1 : 4); ++q) {
      pos[4 * e + q] = ybus_position(*t, rc4[q][0], rc4[q][1]);
      if (pos[4 * e + q] < 0) return fail(GX_ERR_ARG, "branch " + std::to_string(e) + " is not in the Y-bus pattern");
      y[8 * e + 2 * q] = b.y[q].re;
      y[8 * e + 2 * q + 1] = b.y[q].im;
    }
  }
  t->n_br = (int)br.size();
  t->br_pos.swap(pos);
  t->br_y.swap(y);
  t->br_bridge = find_bridges(t->n_bus, br);
  t->br_uploaded = false;
  return GX_OK;
}

int gx_topo_branch_is_bridge(const gx_topo* t, int branch) {
  if (!t) return fail(GX_ERR_ARG, "null");
  if (branch < 0 || branch >= t->n_br) return fail(GX_ERR_ARG, "branch out of range");
  return t->br_bridge[branch] ? 1 : 0;
}

int gx_batch_solve(gx_topo* t, int n_scen, const double* p_spec, const double* q_spec, const double* vm0,
                   const double* va0, double tol, int max_iter, int mode, int flags, double* vm, double* va,
                   int* status, int* iters, double* history, void* stream_ptr) {
  return gx_batch_solve_outages(t, n_scen, p_spec, q_spec, vm0, va0, nullptr, tol, max_iter, mode, flags, vm, va,
                                status, iters, history, stream_ptr);
}

int gx_batch_solve_outages(gx_topo* t, int n_scen, const double* p_spec, const double* q_spec, const double* vm0,
                           const double* va0, const int* outage, double tol, int max_iter, int mode, int flags,
                           double* vm, double* va, int* status, int* iters, double* history, void* stream_ptr) {
  if (!t) return fail(GX_ERR_ARG, "null topology");
  if (n_scen < 0 || max_iter < 0 || !(tol >= 0.0)) return fail(GX_ERR_ARG, "bad n_scen/max_iter/tol");
  if (mode != GX_MODE_FP64 && mode != GX_MODE_MIXED) return fail(GX_ERR_ARG, "unknown mode");
  if (n_scen == 0) return GX_OK;
  if (!p_spec || !q_spec || !vm0 || !va0 || !vm || !va || !status || !iters) return fail(GX_ERR_ARG, "null array");
  BatchIn in;
  in.n_scen = n_scen;
  in.p = p_spec;
  in.q = q_spec;
  in.vm0 = vm0;
  in.va0 = va0;
  in.tol = tol;
  in.max_iter = max_iter;
  in.fp32 = mode == GX_MODE_MIXED;
  in.flags = flags;
  in.dev = (flags & GX_DEVICE_PTRS) != 0;
  in.st = (hipStream_t)stream_ptr;
  std::vector<int> h_outage;
  bool any_outage = false;
  int rc;
  if ((rc = read_outages(*t, n_scen, outage, in.dev, h_outage, any_outage))) return rc;
  in.outage = any_outage ? &h_outage : nullptr;
  hipStream_t st = in.st;
  const int tile = t->tile_req ? t->tile_req : auto_tile(n_scen, in.fp32);
  // 8 scenarios per tile: pair tiles (4-lane groups, two scenarios per lane), whose FP32 refactor
  // runs as quad tiles (2-lane groups, four scenarios per lane) unless 16 refactor groups were
  // asked for
  const bool pairs = tile == 8;
  const bool quads = pairs && in.fp32 && t->refactor_groups != 16;
  gx::DevWork w;
  Workspace& ws = t->ws;
  HostResult hres;
  if ((rc = stage_batch(*t, ws, in, tile, w, &hres, pairs, quads))) return rc;
  // (quad tiles in use: the FP64 re-solve below switches the refactor to the pair groups)
  const bool quad_run = quads && t->dt.groups == gx::kLanes / (tile / 2) && t->dt.part_groups == 2 * t->dt.groups;
  if (hres.used) return host_results(*t, in, hres, vm, va, status, iters, history);
  // optional per-launch timing of each kernel class on this stream
  auto timed = [&](int cls, auto&& launch) -> int {
    if (!t->profiling) {
      launch();
      return GX_OK;
    }
    hipEvent_t ab[2];
    for (hipEvent_t& e : ab) {
      if (t->ev_pool.empty()) {
        GX_HIP(hipEventCreate(&e));
      } else {
        e = t->ev_pool.back();
        t->ev_pool.pop_back();
      }
    }
    GX_HIP(hipEventRecord(ab[0], st));
    launch();
    GX_HIP(hipEventRecord(ab[1], st));
    t->prof.emplace_back(cls, ab[0], ab[1]);
    return GX_OK;
  };
  // the Newton loop (newton_raphson_cached, src/solver.rs:663-736) for every running scenario
  auto newton = [&](bool lu32) -> int {
    for (int it = 0; it < max_iter; ++it) {
      int rc2;
      // (the mismatch kernel also writes the Jacobian when the topology has its Y-position map and
      // the refactor stages structural fill as zero itself; else the fill launch also zeroes those)
      const bool fused = t->dt.y_lupos != nullptr && t->dt.stage_zero;
      if ((rc2 = timed(GX_K_MISMATCH, [&] { gx::launch_mismatch(t->dt, w, lu32, st, fused); }))) return rc2;
      gx::launch_decide(t->dt, w, it, st);
      // stop as soon as no scenario is left running (no empty launches after convergence)
      GX_HIP(hipMemcpyAsync(ws.h_running + it, w.running + it, sizeof(int), hipMemcpyDeviceToHost, st));
      GX_HIP(hipStreamSynchronize(st));
      if (ws.h_running[it] == 0) break;
      if (!fused && (rc2 = timed(GX_K_FILL, [&] { gx::launch_fill(t->dt, w, lu32, st); }))) return rc2;
      if ((rc2 = timed(GX_K_LU, [&] { gx::launch_lu(t->dt, w, lu32, true, it, st); }))) return rc2;
      if (lu32) {
        if ((rc2 = timed(GX_K_REFINE, [&] {
               gx::launch_refine(t->dt, w, ws.dvr.as<double>(), ws.dvi.as<double>(), it, st);
             })))
          return rc2;
        if ((rc2 = timed(GX_K_SOLVE, [&] { gx::launch_lu(t->dt, w, lu32, false, it, st); }))) return rc2;
      }
      if ((rc2 = timed(GX_K_UPDATE, [&] { gx::launch_update(t->dt, w, lu32, lu32, it, st); }))) return rc2;
    }
    gx::launch_finish(w, st);
    GX_HIP(hipGetLastError());
    return GX_OK;
  };
  t->last_fallback = 0;
  if (in.fp32) {
    w.stall_check = 1;
    if ((rc = newton(true))) return rc;
    // per-instance FP64 fallback: every scenario that stalled or met a singular FP32 factor is
    // solved again from its starting point with the FP64 path
    w.stall_check = 0;
    gx::launch_fallback(t->dt, w, st);
    GX_HIP(hipMemcpyAsync(&ws.h_running[0], w.fallback, sizeof(int), hipMemcpyDeviceToHost, st));
    GX_HIP(hipStreamSynchronize(st));
    t->last_fallback = ws.h_running[0];
    if (t->last_fallback > 0) {
      if ((rc = ws.lu.reserve(lu_lanes((size_t)w.n_tiles * w.tile, w.tile) * (size_t)t->dt.nnz_lu * 8))) return rc;
      w.lu = ws.lu.p;
      // the FP64 re-solve of pair tiles refactors on the solves' lane groups
      if (quad_run && (rc = upload_topo(*t, t->dt.window, tile / 2, tile / 2))) return rc;
      if ((rc = newton(false))) return rc;
    }
  } else {
    if ((rc = newton(false))) return rc;
  }
  t->last = w;
  t->has_last = true;
  if (flags & GX_KEEP_WARM) {
    if ((rc = ensure_warm(*t))) return rc;
    gx::launch_keep_warm(t->dt, w, t->d_warm.as<double>(), t->d_warm_ok.as<int>(), st);
  }
  if ((rc = collect_batch(*t, ws, w, in.dev, vm, va, status, iters, history, st))) return rc;
  return fold_profile(*t);  // (the collection waited for the stream)
}

}  // extern "C"

// ---------------------------------------------------------------- LinearSolver-level system
struct gx_sys {
  int n = 0;
  bool on_device = false;
  size_t n_trip = 0;
  gx::CscPattern csc;
  std::vector<int> slot, csc2lu, qinv;
  gx::Ordering ord;
  gx::LuPattern lu;
  LuSchedule sched;
  DevSched ds;
  DevMem d_lu, d_y, d_status, d_iters, d_vals, d_rhs;
  // tables of the allocation-free solve: per L+U position the triplets summed into it (in triplet
  // order), the row order of the right-hand side (prow[k] = original row of pivot step k) and the
  // column order of the solution
  DevMem d_pk_ptr, d_pk_idx, d_prow, d_q;
  gx::DevTopo dt;
  gx::DevWork w;
};

namespace {
int sys_to_device(gx_sys* s);
}

extern "C" {

gx_sys* gx_sys_create(int n, int nnz, const int* rows, const int* cols, const double* vals) {
  if (n <= 0 || nnz <= 0 || !rows || !cols || !vals) {
    fail(GX_ERR_ARG, "bad system");
    return nullptr;
  }
  for (int e = 0; e < nnz; ++e)
    if (rows[e] < 0 || rows[e] >= n || cols[e] < 0 || cols[e] >= n) {
      fail(GX_ERR_ARG, "triplet out of range");
      return nullptr;
    }
  std::unique_ptr<gx_sys> s(new gx_sys());
  s->n = n;
  s->n_trip = (size_t)nnz;
  std::vector<int> r(rows, rows + nnz), c(cols, cols + nnz);
  triplets_to_csc(n, r, c, s->csc, s->slot);
  std::vector<double> ax(s->csc.ai.size(), 0.0);
  for (int e = 0; e < nnz; ++e) ax[s->slot[e]] += vals[e];
  if (!gx::analyze(s->csc, s->ord) || !gx::factor_pivot(s->csc, ax.data(), s->ord, 1e-3, s->lu)) {
    fail(GX_ERR_SINGULAR, "singular system");
    return nullptr;
  }
  if (build_schedule(s->lu, s->sched)) return nullptr;
  s->qinv.assign(n, 0);
  for (int k = 0; k < n; ++k) s->qinv[s->lu.q[k]] = k;
  s->csc2lu.assign(s->csc.ai.size(), -1);
  for (int j = 0; j < n; ++j)
    for (int p = s->csc.ap[j]; p < s->csc.ap[j + 1]; ++p)
      s->csc2lu[p] = lu_position(s->lu, s->lu.pinv[s->csc.ai[p]], s->qinv[j]);
  // with a GPU present the system's device tables are uploaded here, at set-up, so that its solves
  // allocate nothing (a failed upload is retried, and reported, by the first solve)
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0) (void)sys_to_device(s.get());
  else (void)hipGetLastError();
  return s.release();
}

}  // extern "C"

namespace {

// Device resources of a system, allocated at its first solve.
// One system is solved as lane 0 of a single tile of kSysTile scenarios.
constexpr int kSysTile = 4;

int sys_to_device(gx_sys* s) {
  if (s->on_device) return GX_OK;
  const int n = s->n, nlu = s->lu.nnz();
  const int rows = pick_ring_rows(lu_cap(s->sched.max_col, kSysTile), kSysTile, sizeof(double), 0);
  if (rows < 0) return fail(GX_ERR_UNSUPPORTED, "longest L+U column exceeds the LDS staging limit");
  int rc;
  if ((rc = upload_sched(s->lu, s->sched, gx::ring_window_of(rows, kSysTile), kSysTile, s->ds, s->dt)) ||
      (rc = s->d_lu.reserve((size_t)nlu * lu_lanes(kSysTile, kSysTile) * sizeof(double))) ||
      (rc = s->d_y.reserve((size_t)n * kSysTile * sizeof(double))) ||
      (rc = s->d_status.reserve(kSysTile * sizeof(int))) || (rc = s->d_iters.reserve(kSysTile * sizeof(int))) ||
      (rc = s->d_vals.reserve((size_t)nlu * sizeof(double))) || (rc = s->d_rhs.reserve((size_t)n * sizeof(double))))
    return rc;
  GX_HIP(hipMemset(s->d_lu.p, 0, s->d_lu.bytes));
  GX_HIP(hipMemset(s->d_y.p, 0, s->d_y.bytes));
  {
    std::vector<int> ptr(nlu + 1, 0), idx(s->n_trip), prow(n), q(s->lu.q.begin(), s->lu.q.end());
    for (size_t e = 0; e < s->n_trip; ++e) ++ptr[s->csc2lu[s->slot[e]] + 1];
    for (int p = 0; p < nlu; ++p) ptr[p + 1] += ptr[p];
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (size_t e = 0; e < s->n_trip; ++e) idx[fill[s->csc2lu[s->slot[e]]]++] = (int)e;
    for (int r = 0; r < n; ++r) prow[s->lu.pinv[r]] = r;
    if ((rc = s->d_pk_ptr.upload(ptr)) || (rc = s->d_pk_idx.upload(idx)) || (rc = s->d_prow.upload(prow)) ||
        (rc = s->d_q.upload(q)))
      return rc;
  }
  s->w.tile = kSysTile;
  s->w.n_scen = 1;
  s->w.n_tiles = 1;
  s->w.lu = s->d_lu.p;
  s->w.y = s->d_y.p;
  s->w.status = s->d_status.as<int>();
  s->w.iters = s->d_iters.as<int>();
  s->on_device = true;
  return GX_OK;
}

}  // namespace

extern "C" {

int gx_sys_ordering(const gx_sys* s, int* p, int* q, int* r, int* pnum, int* nblocks) {
  if (!s) return fail(GX_ERR_ARG, "null");
  if (p) std::memcpy(p, s->ord.p.data(), sizeof(int) * s->n);
  if (q) std::memcpy(q, s->ord.q.data(), sizeof(int) * s->n);
  if (r) std::memcpy(r, s->ord.r.data(), sizeof(int) * s->ord.r.size());
  if (pnum) std::memcpy(pnum, s->lu.pnum.data(), sizeof(int) * s->n);
  if (nblocks) *nblocks = (int)s->ord.r.size() - 1;
  return s->lu.nnz();
}

int gx_sys_factor_and_solve_values(gx_sys* s, const double* vals, const double* rhs, double* x) {
  if (!s || !vals || !rhs || !x) return fail(GX_ERR_ARG, "null");
  int rc0 = sys_to_device(s);
  if (rc0) return rc0;
  const int n = s->n, nlu = s->lu.nnz();
  std::vector<double> packed(nlu, 0.0), b(n), out(n);
  for (size_t e = 0; e < s->n_trip; ++e) packed[s->csc2lu[s->slot[e]]] += vals[e];
  for (int r = 0; r < n; ++r) b[s->lu.pinv[r]] = rhs[r];
  std::vector<int> st(kSysTile, gx::kPadLane);
  st[0] = gx::kRunning;
  GX_HIP(hipMemcpy(s->d_vals.p, packed.data(), sizeof(double) * nlu, hipMemcpyHostToDevice));
  GX_HIP(hipMemcpy(s->d_rhs.p, b.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  GX_HIP(hipMemcpy(s->d_status.p, st.data(), sizeof(int) * kSysTile, hipMemcpyHostToDevice));
  // lane 0 of tile 0: consecutive positions lie a block of tiles' values apart
  gx::launch_lane0(s->d_vals.as<double>(), s->d_lu.as<double>(), nlu, true, kSysTile * gx::lu_block(kSysTile), nullptr);
  gx::launch_lane0(s->d_rhs.as<double>(), s->d_y.as<double>(), n, true, kSysTile, nullptr);
  gx::launch_lu(s->dt, s->w, false, true, 0, nullptr);
  gx::launch_lane0(s->d_y.as<double>(), s->d_rhs.as<double>(), n, false, kSysTile, nullptr);
  GX_HIP(hipGetLastError());
  GX_HIP(hipMemcpy(b.data(), s->d_rhs.p, sizeof(double) * n, hipMemcpyDeviceToHost));
  GX_HIP(hipMemcpy(st.data(), s->d_status.p, sizeof(int), hipMemcpyDeviceToHost));
  if (st[0] != gx::kRunning) return GX_SINGULAR;
  for (int k = 0; k < n; ++k) x[s->lu.q[k]] = b[k];
  return GX_OK;
}

int gx_sys_factor_and_solve_values_noalloc(gx_sys* s, const double* vals, const double* rhs, double* x,
                                           void* stream) {
  if (!s || !vals || !rhs || !x) return fail(GX_ERR_ARG, "null");
  static_assert(gx::kSysTileW == kSysTile, "one tile layout for both system solves");
  if (!s->on_device) return fail(GX_ERR_STATE, "the system's device tables are not resident (gx_sys_create saw no GPU)");
  const int n = s->n, nlu = s->lu.nnz();
  auto up256 = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t lu_b = up256((size_t)nlu * lu_lanes(kSysTile, kSysTile) * sizeof(double));
  const size_t y_b = up256((size_t)n * kSysTile * sizeof(double));
  const size_t need = lu_b + y_b + 256;
  if (need > gx::kScratchBytes)
    return fail(GX_ERR_CAPACITY, "system needs " + std::to_string(need) + " bytes of device scratch; the allocation-free "
                                 "solve holds " + std::to_string(gx::kScratchBytes) +
                                 " (" + std::to_string(lu_lanes(kSysTile, kSysTile) * sizeof(double)) +
                                 " B per L+U entry + 32 B per row): use gx_sys_factor_and_solve_values");
  void* base = nullptr;
  GX_HIP(gx::scratch_base(&base));
  char* b = static_cast<char*>(base);
  gx::DevWork w = s->w;
  w.lu = b;
  w.y = b + lu_b;
  w.status = reinterpret_cast<int*>(b + lu_b + y_b);
  w.iters = w.status + kSysTile;
  hipStream_t st = (hipStream_t)stream;
  gx::launch_sys_stage(s->d_pk_ptr.as<int>(), s->d_pk_idx.as<int>(), vals, nlu, s->d_prow.as<int>(), rhs, n,
                       static_cast<double*>(w.lu), static_cast<double*>(w.y), w.status, w.iters, st);
  gx::launch_lu(s->dt, w, false, true, 0, st);
  gx::launch_sys_unpack(static_cast<const double*>(w.y), s->d_q.as<int>(), n, x, st);
  GX_HIP(hipGetLastError());
  int st0 = gx::kSingular;
  GX_HIP(hipMemcpyAsync(&st0, w.status, sizeof(int), hipMemcpyDeviceToHost, st));
  GX_HIP(hipStreamSynchronize(st));
  return st0 == gx::kRunning ? GX_OK : GX_SINGULAR;
}

void gx_sys_destroy(gx_sys* s) { delete s; }

}  // extern "C"

// ---------------------------------------------------------------- request engine
// Many batches ("requests": one topology, its scenarios, tolerance, iteration limit, mode) in
// flight at once.  Each step runs one Newton iteration of every active request: its element
// kernels, then ONE persistent LU launch over the tiles of all requests (heaviest tiles first,
// pulled from a device work queue; tiles whose scenarios all stopped return at once), then its
// update.  Requests join between steps (admitted while their workspace fits the device-memory
// budget) and leave when finished, so new requests overlap on the device with those already
// running.  gx_engine_run captures one step of the current request set in a hipGraph and replays it.
struct gx_request {
  int id = 0;
  gx_topo* t = nullptr;
  Workspace ws;
  gx::DevWork w;
  bool dev = false;  // results go to device pointers
  bool mixed = false, lu32 = false;
  int iter = 0;   // Newton iteration of the current phase (host mirror of w.iter_dev)
  int phase = 0;  // 0 first solve, 1 FP64 re-solve of the mixed mode's fallbacks, 2 finished
  int fallbacks = 0;
  int flags = 0;
  double cost = 0.0;  // relative LU time of one tile (queue order)
  long long reserved = 0;  // bytes admitted against the budget (the estimate at submission)
};

struct gx_engine {
  int tile = 4;
  int flags = 0;
  long long budget = 0;
  hipStream_t st = nullptr;
  bool own_stream = false;
  int n_cu = 256;
  int next_id = 0;
  std::map<int, std::unique_ptr<gx_request>> reqs;
  std::vector<gx_request*> active;
  bool dirty = true;  // LU tables / captured step out of date
  // LU launch kinds: 0 FP64 factor+solve, 1 FP32 factor+solve, 2 FP32 correction solve
  struct Launch {
    DevMem topos, works, items, counter;
    int n_items = 0;
    size_t shmem = 0;
    int grid = 0;
  } lk[3];
  int* h_run = nullptr;  // pinned: running count of each active request after its last decision
  size_t h_run_cap = 0;
  hipGraphExec_t exec = nullptr;
  ~gx_engine() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (h_run) (void)hipHostFree(h_run);
    reqs.clear();
    if (own_stream && st) (void)hipStreamDestroy(st);
  }
};

namespace {

// Rebuilds the device tables of the three LU launch kinds for the active set (synchronous: only
// when the set or a request's phase changed).
int engine_tables(gx_engine& e) {
  for (int k = 0; k < 3; ++k) {
    gx_engine::Launch& L = e.lk[k];
    std::vector<gx::DevTopo> topos;
    std::vector<gx::DevWork> works;
    std::vector<std::pair<double, int2>> items;
    L.shmem = 0;
    for (gx_request* r : e.active) {
      const bool fp32 = r->lu32;
      if ((k == 0 && fp32) || (k > 0 && !fp32)) continue;
      const int idx = (int)topos.size();
      if (e.tile == 8 && r->t->dt.groups == gx::kLanes / 4) {
        // pair tiles: the refactor parts of this launch's precision (quad tiles for FP32 where the
        // 32-group parts exist; the topology's schedule caches both part sets)
        int rc;
        const int window = r->t->dt.window;
        if ((rc = upload_topo(*r->t, window, 4, fp32 ? 2 : 4))) return rc;
        if (fp32 && r->t->dt.part_groups != gx::kLanes / 2 && (rc = upload_topo(*r->t, window, 4, 4))) return rc;
      }
      topos.push_back(r->t->dt);
      works.push_back(r->w);
      for (int tl = 0; tl < r->w.n_tiles; ++tl) items.push_back({r->cost, make_int2(idx, tl)});
      L.shmem = std::max(L.shmem, gx::lu_shmem_tile(r->t->dt, fp32, k < 2, e.tile));
    }
    std::stable_sort(items.begin(), items.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
    std::vector<int2> it2(items.size());
    for (size_t i = 0; i < items.size(); ++i) it2[i] = items[i].second;
    L.n_items = (int)it2.size();
    if (L.n_items == 0) continue;
    int rc;
    if ((rc = L.topos.reserve(sizeof(gx::DevTopo) * topos.size())) ||
        (rc = L.works.reserve(sizeof(gx::DevWork) * works.size())) || (rc = L.items.reserve(sizeof(int2) * it2.size())) ||
        (rc = L.counter.reserve(sizeof(int))))
      return rc;
    GX_HIP(hipMemcpyAsync(L.topos.p, topos.data(), sizeof(gx::DevTopo) * topos.size(), hipMemcpyHostToDevice, e.st));
    GX_HIP(hipMemcpyAsync(L.works.p, works.data(), sizeof(gx::DevWork) * works.size(), hipMemcpyHostToDevice, e.st));
    GX_HIP(hipMemcpyAsync(L.items.p, it2.data(), sizeof(int2) * it2.size(), hipMemcpyHostToDevice, e.st));
    GX_HIP(hipStreamSynchronize(e.st));
    // resident wavefronts: LDS and (at most) eight per CU
    const int per_cu = std::max(1, std::min(8, (int)(gx::kMaxLds / std::max<size_t>(L.shmem, 1))));
    L.grid = std::min(L.n_items, e.n_cu * per_cu);
  }
  if (e.active.size() > e.h_run_cap) {
    if (e.h_run) (void)hipHostFree(e.h_run);
    e.h_run = nullptr;
    GX_HIP(hipHostMalloc((void**)&e.h_run, sizeof(int) * e.active.size(), hipHostMallocDefault));
    e.h_run_cap = e.active.size();
  }
  e.dirty = false;
  if (e.exec) {
    (void)hipGraphExecDestroy(e.exec);
    e.exec = nullptr;
  }
  return GX_OK;
}

// Launches of one Newton step of every active request (no host synchronization inside).
void engine_step_launches(gx_engine& e) {
  // (the mismatch kernel also writes the Jacobian where the topology allows it, as in a batch solve)
  auto fused = [](const gx_request* r) { return r->t->dt.y_lupos != nullptr && r->t->dt.stage_zero; };
  for (gx_request* r : e.active) {
    gx::launch_mismatch(r->t->dt, r->w, r->lu32, e.st, fused(r));
    gx::launch_decide(r->t->dt, r->w, 0, e.st);
  }
  for (gx_request* r : e.active)
    if (!fused(r)) gx::launch_fill(r->t->dt, r->w, r->lu32, e.st);
  for (int k = 0; k < 2; ++k) {
    gx_engine::Launch& L = e.lk[k];
    gx::launch_lu_queue(e.tile, L.topos.as<gx::DevTopo>(), L.works.as<gx::DevWork>(), L.items.as<int2>(), L.n_items,
                        L.counter.as<int>(), L.shmem, L.grid, k == 1, true, e.st);
  }
  for (gx_request* r : e.active)
    if (r->lu32) gx::launch_refine(r->t->dt, r->w, r->ws.dvr.as<double>(), r->ws.dvi.as<double>(), 0, e.st);
  {
    gx_engine::Launch& L = e.lk[2];
    gx::launch_lu_queue(e.tile, L.topos.as<gx::DevTopo>(), L.works.as<gx::DevWork>(), L.items.as<int2>(), L.n_items,
                        L.counter.as<int>(), L.shmem, L.grid, true, false, e.st);
  }
  for (gx_request* r : e.active) {
    gx::launch_update(r->t->dt, r->w, r->lu32, r->lu32, 0, e.st);
    gx::launch_iter_advance(r->w.iter_dev, e.st);
  }
}

// A request whose current Newton phase ended (no scenario running, or its iteration limit):
// MaxIterationsReached for those still running; in the mixed mode, the stalled or FP32-singular
// scenarios start the FP64 re-solve; otherwise the request is finished.
int engine_end_phase(gx_engine& e, gx_request& r) {
  gx::launch_finish(r.w, e.st);
  if (r.mixed && r.phase == 0) {
    r.w.stall_check = 0;
    gx::launch_fallback(r.t->dt, r.w, e.st);
    GX_HIP(hipMemcpyAsync(&r.ws.h_running[0], r.w.fallback, sizeof(int), hipMemcpyDeviceToHost, e.st));
    GX_HIP(hipStreamSynchronize(e.st));
    r.fallbacks = r.ws.h_running[0];
    if (r.fallbacks > 0) {
      // the FP64 re-solve needs FP64-sized L+U values (the stream is idle here)
      int rc;
      if ((rc = r.ws.lu.reserve(lu_lanes((size_t)r.w.n_tiles * r.w.tile, r.w.tile) * (size_t)r.t->dt.nnz_lu * 8)))
        return rc;
      r.w.lu = r.ws.lu.p;
      r.phase = 1;
      r.lu32 = false;
      r.iter = 0;
      e.dirty = true;
      return GX_OK;
    }
  }
  r.phase = 2;
  e.dirty = true;
  if (r.flags & GX_KEEP_WARM) {
    int rc;
    if ((rc = ensure_warm(*r.t))) return rc;
    gx::launch_keep_warm(r.t->dt, r.w, r.t->d_warm.as<double>(), r.t->d_warm_ok.as<int>(), e.st);
  }
  return GX_OK;
}

// After steps: each request's running count after its last decision; phase ends.
int engine_retire(gx_engine& e) {
  for (size_t i = 0; i < e.active.size(); ++i) {
    gx_request* r = e.active[i];
    GX_HIP(hipMemcpyAsync(e.h_run + i, r->w.running + (r->iter - 1), sizeof(int), hipMemcpyDeviceToHost, e.st));
  }
  GX_HIP(hipStreamSynchronize(e.st));
  std::vector<gx_request*> keep;
  for (size_t i = 0; i < e.active.size(); ++i) {
    gx_request* r = e.active[i];
    int rc;
    if (e.h_run[i] == 0 || r->iter >= r->w.max_iter) {
      if ((rc = engine_end_phase(e, *r))) return rc;
    }
    if (r->phase < 2) keep.push_back(r);
  }
  if (keep.size() != e.active.size()) e.dirty = true;
  e.active.swap(keep);
  return GX_OK;
}

}  // namespace

extern "C" {

gx_engine* gx_engine_create(int tile, int flags, long long hbm_budget, void* stream, int* err) {
  auto set = [&](int c) {
    if (err) *err = c;
  };
  if (tile != 0 && !valid_tile(tile)) {
    set(fail(GX_ERR_ARG, "tile must be 0, 4, 8, 16 or 64"));
    return nullptr;
  }
  std::unique_ptr<gx_engine> e(new gx_engine());
  e->tile = tile ? tile : 4;
  e->flags = flags;
  e->budget = hbm_budget;
  e->st = (hipStream_t)stream;
  if (!e->st) {
    if (hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking) != hipSuccess) {
      set(fail(GX_ERR_DEVICE, "hipStreamCreate"));
      return nullptr;
    }
    e->own_stream = true;
  }
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
      cus > 0)
    e->n_cu = cus;
  set(GX_OK);
  return e.release();
}

void gx_engine_destroy(gx_engine* e) {
  if (e && e->st) (void)hipStreamSynchronize(e->st);
  delete e;
}

long long gx_engine_bytes(const gx_engine* e) {
  if (!e) return fail(GX_ERR_ARG, "null");
  long long b = 0;
  for (auto& kv : e->reqs) b += std::max((long long)kv.second->ws.bytes(), kv.second->reserved);
  return b;
}

long long gx_engine_request_bytes(const gx_topo* t, int n_scen, int max_iter, int mode, int flags, int tile) {
  if (!t || n_scen <= 0 || max_iter < 0) return fail(GX_ERR_ARG, "bad request");
  tile = tile ? tile : 4;
  const long long lanes = (long long)((n_scen + tile - 1) / tile) * tile;
  const long long nb = t->n_bus, n = t->n;
  const long long nnz_lu = t->factored ? t->lu.nnz() : 4 * (long long)t->jkind.size();
  long long b = lanes * (8 * nb * 8 + 3 * n * 8 + (long long)std::max(max_iter, 1) * 8 + 16) +
                (long long)lu_lanes((size_t)lanes, tile) * nnz_lu * (mode == GX_MODE_MIXED ? 4 : 8) +
                (long long)n_scen * nb * 8 + 4LL * std::max(max_iter, 1) + 64;
  if (mode == GX_MODE_MIXED) b += lanes * 4 * nb * 8;
  if (flags & GX_INIT_LINEAR) b += lanes * (4 * nnz_lu / 2 * 16 + nb * 16 + 8);
  return b;
}

int gx_engine_submit(gx_engine* e, gx_topo* t, int n_scen, const double* p_spec, const double* q_spec,
                     const double* vm0, const double* va0, const int* outage, double tol, int max_iter, int mode,
                     int flags) {
  if (!e || !t) return fail(GX_ERR_ARG, "null engine or topology");
  if (n_scen <= 0 || max_iter <= 0 || !(tol >= 0.0)) return fail(GX_ERR_ARG, "bad n_scen/max_iter/tol");
  if (mode != GX_MODE_FP64 && mode != GX_MODE_MIXED) return fail(GX_ERR_ARG, "unknown mode");
  if (!p_spec || !q_spec || !vm0 || !va0) return fail(GX_ERR_ARG, "null array");
  const long long need = gx_engine_request_bytes(t, n_scen, max_iter, mode, flags, e->tile);
  if (e->budget > 0 && gx_engine_bytes(e) + need > e->budget)
    return fail(GX_ERR_BUDGET, "request does not fit the device-memory budget now");
  std::unique_ptr<gx_request> r(new gx_request());
  r->reserved = need;
  BatchIn in;
  in.n_scen = n_scen;
  in.p = p_spec;
  in.q = q_spec;
  in.vm0 = vm0;
  in.va0 = va0;
  in.tol = tol;
  in.max_iter = max_iter;
  in.fp32 = mode == GX_MODE_MIXED;
  in.flags = flags;
  in.dev = (flags & GX_DEVICE_PTRS) != 0;
  in.st = e->st;
  std::vector<int> h_outage;
  bool any = false;
  int rc;
  if ((rc = read_outages(*t, n_scen, outage, in.dev, h_outage, any))) return rc;
  in.outage = any ? &h_outage : nullptr;
  // (at 8 scenarios per tile, pair tiles with quad-tile FP32 refactors, as in gx_batch_solve)
  const bool pairs = e->tile == 8;
  if ((rc = stage_batch(*t, r->ws, in, e->tile, r->w, nullptr, pairs, pairs && in.fp32))) return rc;
  if ((rc = r->ws.iter_dev.reserve(sizeof(int)))) return rc;
  r->w.iter_dev = r->ws.iter_dev.as<int>();
  gx::launch_zero_ints(r->w.iter_dev, 1, e->st);
  r->id = e->next_id++;
  r->t = t;
  r->dev = in.dev;
  r->mixed = in.fp32;
  r->lu32 = in.fp32;
  r->w.stall_check = in.fp32 ? 1 : 0;
  r->flags = flags;
  r->cost = (double)t->dt.n * (1.0 + (double)t->sched.n_upd / std::max(1, t->dt.n) / 16.0);
  GX_HIP(hipGetLastError());
  const int id = r->id;
  e->active.push_back(r.get());
  e->reqs[id] = std::move(r);
  e->dirty = true;
  return id;
}

int gx_engine_step(gx_engine* e) {
  if (!e) return fail(GX_ERR_ARG, "null");
  if (e->active.empty()) return 0;
  int rc;
  if (e->dirty && (rc = engine_tables(*e))) return rc;
  engine_step_launches(*e);
  for (gx_request* r : e->active) ++r->iter;
  GX_HIP(hipGetLastError());
  if ((rc = engine_retire(*e))) return rc;
  return (int)e->active.size();
}

int gx_engine_run(gx_engine* e, int max_steps) {
  if (!e || max_steps < 0) return fail(GX_ERR_ARG, "bad arguments");
  int rc;
  while (!e->active.empty() && max_steps > 0) {
    if (e->dirty && (rc = engine_tables(*e))) return rc;
    // replay as many steps as no request can finish its phase by its iteration limit
    int n = max_steps;
    for (gx_request* r : e->active) n = std::min(n, r->w.max_iter - r->iter);
    n = std::max(n, 1);
    if (!e->exec) {
      hipGraph_t g = nullptr;
      GX_HIP(hipStreamBeginCapture(e->st, hipStreamCaptureModeThreadLocal));
      engine_step_launches(*e);
      GX_HIP(hipStreamEndCapture(e->st, &g));
      const hipError_t ie = hipGraphInstantiate(&e->exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) return fail(GX_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    }
    // check for converged requests every few steps
    const int chunk = std::min(n, 4);
    for (int k = 0; k < chunk; ++k) GX_HIP(hipGraphLaunch(e->exec, e->st));
    for (gx_request* r : e->active) r->iter += chunk;
    max_steps -= chunk;
    if ((rc = engine_retire(*e))) return rc;
  }
  return (int)e->active.size();
}

int gx_engine_collect(gx_engine* e, int id, double* vm, double* va, int* status, int* iters, double* history) {
  if (!e) return fail(GX_ERR_ARG, "null");
  auto it = e->reqs.find(id);
  if (it == e->reqs.end()) return fail(GX_ERR_ARG, "unknown request id");
  gx_request& r = *it->second;
  if (r.phase < 2) return 0;
  if (!vm || !va || !status || !iters) return fail(GX_ERR_ARG, "null output");
  int rc;
  if ((rc = collect_batch(*r.t, r.ws, r.w, r.dev, vm, va, status, iters, history, e->st))) return rc;
  e->reqs.erase(it);
  return 1;
}

int gx_engine_fallbacks(const gx_engine* e, int id) {
  if (!e) return fail(GX_ERR_ARG, "null");
  auto it = e->reqs.find(id);
  if (it == e->reqs.end()) return fail(GX_ERR_ARG, "unknown request id");
  return it->second->fallbacks;
}

}  // extern "C"
