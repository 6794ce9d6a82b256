// Device-resident topology and per-batch workspace shared by the host driver and the HIP kernels.
//
// Layout: every per-scenario array is "scenario-minor interleaved" in tiles of TW scenarios
// (TW = DevWork::tile, 4, 8 or 16): element e of scenario s lives at
//   base[(tile * count + e) * TW + s % TW],   tile = s / TW,
// so the same element of a tile's scenarios is one contiguous segment.
// The topology arrays (Y-bus, Jacobian recipes, LU pattern and update schedule) are shared
// read-only by every scenario.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace gx {

constexpr int kLanes = 64;        // wavefront width
constexpr int kMaxTile = 16;      // widest narrow tile (scenarios per tile): 2, 4, 8 or 16 are compiled
constexpr int kRingGap = 32;      // solve ring: columns between fences / rows kept behind the front
constexpr int kChunkRounds = 16;  // refactor update rounds fetched into registers at a time
constexpr int kPartTail = 320;   // refactor part: descriptors of entries past the registered ones in LDS
constexpr int kPartRegs = 4;     // refactor part entries per lane group staged through registers
constexpr int kPartMinChunks = 2;    // round chunks per refactor part at least (= the value-load distance)
constexpr int kPartAhead = 4;        // chunks between loading a chunk's item positions and running it
constexpr int kPartMaxEntries = 1023;  // refactor part entries at most (10-bit offsets in item words)
constexpr int kStageZero = 1 << 30;    // pent_pos flag: the entry is structural fill, staged as zero
constexpr int kJointRounds = 1;        // refactor rounds whose LDS reads are issued before their writes
// Update rounds per chunk of the refactor-by-parts round stream for G lane groups: 8 at 32 groups
// (a quad-tile lane holds four FP32 scenarios, so a chunk's L values take 4 VGPRs per round and the
// three chunks in flight would otherwise need 192), else 16.
constexpr int part_chunk_rounds(int groups) { return groups >= 32 ? 8 : 16; }
// Row stride (values) of the refactor's part buffers for lane groups of tw lanes holding v scenarios
// each: one value of padding per row, none for four-scenario lanes (16-byte values: the buffers of
// a 384-entry part then fit four wavefronts per CU).
constexpr int part_row_stride(int tw, int v) { return v >= 4 ? tw : tw + 1; }
// tail descriptors per part held in LDS for staging capacity cap: the entries past the registered
// ones, at most kPartTail
constexpr int part_tail(int cap, int tile) {
  return cap - (64 / tile) * kPartRegs < 0 ? 0 : (cap - (64 / tile) * kPartRegs > kPartTail ? kPartTail : cap - (64 / tile) * kPartRegs);
}
constexpr int kMaxLds = 160 * 1024;  // LDS bytes per workgroup on gfx950

constexpr int kWideTile = 64;     // one lane per scenario: every index of the LU kernel is uniform
// tiles whose L+U values interleave position by position.  One: each tile's values are a region of
// their own, a position's TW values contiguous, so the L column a refactor update reads (consecutive
// positions) fills whole cache lines of that tile's values instead of sharing each line with tiles
// that run at other times (10,000-bus grid, 8,192 scenarios, FP32: LU launch 45.4 -> 38.4 ms and
// 282 -> 210 GB fetched per launch against blocks of two tiles)
constexpr int lu_block(int tile) { (void)tile; return 1; }
constexpr int kRoundItems = 16;   // update items per round of the wide-tile LU kernel
constexpr int kWideStage = 16;    // entries of a column the wide-tile LU kernel prefetches
constexpr int kWideFar = 4;       // far entries of an entering row the wide-tile LU kernel prefetches

// Default rows of the LDS ring that holds the moving window of the right-hand side during the
// triangular solves, for a tile width (a power of two).
constexpr int ring_rows(int tile) { return tile == kWideTile ? 256 : (tile <= 8 ? 512 : 256); }
// Pushes of a triangular solve that reach farther than this many rows are gathered instead.  The
// wide tile keeps no rows behind the front (a lane only reads its own scenario's values).
constexpr int ring_window_of(int rows, int tile) { return tile == kWideTile ? rows - 1 : rows - kRingGap - 1; }
// LDS bytes of the LU kernel for a tile width: the staged column(s) of max_col + 1 rows (factoring
// only; two of max_col + 2 rows for the narrow tiles, whose last row holds the constant -1) and the
// solve ring, each row holding tile (+1 padding on narrow tiles) values of elem bytes; narrow tiles
// factoring also hold two parts' tail descriptors.
constexpr size_t lu_lds_bytes(int max_col, int rows, int tile, size_t elem, bool factor) {
  return tile == kWideTile
             ? ((factor ? (size_t)(max_col < 1 ? 1 : max_col) + 1 : 0) + (size_t)rows) * (size_t)tile * elem
             : ((factor ? 2 * ((size_t)(max_col < 1 ? 1 : max_col) + 2) : 0) + (size_t)rows) * (size_t)(tile + 1) * elem +
                   (factor ? 8 + 2 * (size_t)part_tail(max_col < 1 ? 1 : max_col, tile) * 8 : 0);
}

// The refactor-by-parts kernel overlays its two phases instead: the part buffers (two of max_col + 2
// rows) and the tail descriptors (two ints each, two parts' worth) while factoring, then the solve
// ring from offset 0.
constexpr size_t lu_parts_lds_bytes(int max_col, int rows, int tile, size_t elem, bool factor) {
  const size_t cap = (size_t)(max_col < 1 ? 1 : max_col), row = (size_t)(tile + 1) * elem;
  const size_t parts = factor ? (2 * (cap + 2) * row + 15) / 16 * 16 + 2 * (size_t)part_tail((int)cap, tile) * 8 : 0;
  const size_t ring = (size_t)rows * row;
  return parts > ring ? parts : ring;
}

// Quad tiles: the refactor runs lane groups of part_tile lanes on part_elem-byte values (rows of
// part_row_stride values), the substitutions lane groups of tile lanes on elem-byte values (ring
// rows of tile + 1 values); the phases overlay as above.
constexpr size_t lu_quad_lds_bytes(int max_col, int rows, int tile, size_t elem, int part_tile, size_t part_elem,
                                   bool factor) {
  const size_t cap = (size_t)(max_col < 1 ? 1 : max_col);
  const size_t prow = (size_t)part_row_stride(part_tile, 4) * part_elem;
  const size_t parts = factor ? (2 * (cap + 2) * prow + 15) / 16 * 16 + 2 * (size_t)part_tail((int)cap, part_tile) * 8 : 0;
  const size_t ring = (size_t)rows * (size_t)(tile + 1) * elem;
  return parts > ring ? parts : ring;
}

// Allocation-free system solve (gx_sys_factor_and_solve_values_noalloc): one system is lane 0 of a
// tile of kSysTileW scenarios; its packed L+U values, right-hand side and tile status live in a
// static device array of kScratchBytes per GPU.
constexpr int kSysTileW = 4;
constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;

enum Status : int { kConverged = 0, kMaxIter = 1, kSingular = 2, kRunning = 3, kPadLane = 4, kStalled = 5 };

// device bus codes: the reference's BusType (0 Slack, 1 PV, 2 PQ) plus kDeadBus for a member of
// an island without a Slack bus (a Slack bus pinned to V = 0 with no injection)
constexpr int kDeadBus = 3;

// Jacobian entry kinds, in the reference's numbering (src/jacobian.rs:46-57); kFill marks an
// L+U position that is not a Jacobian entry (structural fill, value 0 before elimination).
enum Kind : int { kHii = 0, kNii = 1, kHik = 2, kNik = 3, kMii = 4, kLii = 5, kMik = 6, kLik = 7, kFill = 8 };

struct DevTopo {
  int n_bus = 0;   // buses
  int n = 0;       // unknowns (rows/cols of the Jacobian)
  int nnz_lu = 0;  // packed L+U entries (column k: U rows, diagonal, L rows)
  int max_col = 0; // longest packed column
  // Y-bus, CSR, rows sorted by column
  const int* y_rowptr = nullptr;
  const int* y_col = nullptr;
  const double* y_g = nullptr;
  const double* y_b = nullptr;
  // per bus: unknown position in pivot-step space (-1 if none)
  const int* rpos_p = nullptr;  // P-mismatch equation row -> pivot row
  const int* rpos_q = nullptr;  // Q-mismatch equation row -> pivot row
  const int* xpos_a = nullptr;  // angle unknown -> pivot column
  const int* xpos_m = nullptr;  // magnitude unknown -> pivot column
  const int* bus_type = nullptr;
  // per bus, 6 doubles: the (P, Q) of the constant-power, constant-current and constant-impedance
  // ZIP terms at |V| = 1; null when no bus has any
  const double* zip = nullptr;
  // per packed L+U position: Jacobian recipe (kind, bus i, bus k, g, b)
  const int* lu_kind = nullptr;
  const int* lu_bi = nullptr;
  const int* lu_bk = nullptr;
  const double* lu_g = nullptr;
  const double* lu_b = nullptr;
  const int* lu_yp = nullptr;   // per packed position: the Y-bus position its recipe reads (-1: fill)
  // the same recipes grouped by bus pair for the fill kernel: for bus i, entries [fb_ptr[i],
  // fb_ptr[i+1]) = its Y-bus neighbours k (k = i included) with the positions of H, N, M, L (i, k)
  // (-1 where the bus types leave one out), the admittance and its Y-bus position; fz_pos lists the
  // positions that are no Jacobian entry (structural fill, zero before each refactor)
  const int* fb_ptr = nullptr;
  const int* fb_k = nullptr;
  const int4* fb_pos = nullptr;
  const double* fb_g = nullptr;
  const double* fb_b = nullptr;
  const int* fb_yp = nullptr;
  // the same four positions by Y-bus position (-1 where absent): the fill done by the mismatch kernel
  const int4* y_lupos = nullptr;
  const int* fz_pos = nullptr;
  int n_fz = 0;
  // branches (per-scenario outages): Y-bus positions of the ff, ft, tf, tt stamps (-1: none) and
  // their values (re, im per stamp)
  const int4* br_pos = nullptr;
  const double* br_y = nullptr;
  // LU pattern
  const int* cp = nullptr;      // n+1
  const int* ulen = nullptr;    // n
  const int* lu_row = nullptr;  // nnz_lu, pivot-step row of each packed entry
  // refactor updates, scheduled for G = 64 / tile lane groups: the updates of column k
  // (x_i -= L(i,j) * U(j,k)) in rounds[k] rounds of G slots, one per lane group, stored in chunks of
  // kChunkRounds rounds; every column owns at least one chunk.  Slot: {x, word}, word = target
  // offset (bits 0-15) | source U offset (bits 16-30) | bit 31 set when L(i,j) is read from the
  // previous column's LDS copy at offset x (else x is its packed position); an empty slot is
  // {0, word aimed at the unused row just past the column}
  const int* chunk0 = nullptr;  // n + 1: first chunk of column k
  const int* rounds = nullptr;  // n
  const int2* slots = nullptr;  // (chunk0[n] + 2) * kChunkRounds * G (two empty chunks at the end)
  int groups = 0;               // G the rounds were built for
  // per column k, for the window and G above:
  //   col_a = {cp[k], column length, ulen[k], chunk0[k]}
  //   col_f = {rounds[k], range [q0, q1) of fo_pos/fo_col: the far L entries of the row
  //            r = k + W + 1 entering the forward window after column k that fslot does not hold,
  //            1 if that row has far L entries at all}
  //   col_b = {range [q0, q1) of bwd_pos/bwd_col: the far U entries of the row r = k - W - 1
  //            entering the backward window after column k that bslot does not hold, 1 if that
  //            row has far U entries at all, 0}
  const int4* col_a = nullptr;
  const int4* col_f = nullptr;
  const int4* col_b = nullptr;
  // far L entries of the row entering the forward window after column k: fslot[k * G + i] = the
  // i-th entry {packed position, column} for i < G ({-1, -1} if none); fo_pos/fo_col the others
  const int2* fslot = nullptr;
  const int* fo_pos = nullptr;
  const int* fo_col = nullptr;
  // wide tile: the updates of column k in rounds [uround[k], uround[k+1]) of kRoundItems items
  // {target offset | source U offset << 16, packed position of L(i,j)} (position -1: empty item
  // aimed at the unused row just past the column); items of a round have distinct targets and
  // read no target of the same round; every column owns at least one round, and the first round
  // of column k reads no L(:,k-1)
  const int* uround = nullptr;
  const int2* uitems = nullptr;
  // first far entries {position, column} of the row entering the backward window after column k
  // ({-1, -1} if none): narrow tiles G per column (entry i for lane group i), the wide tile
  // kWideFar (and as many forward ones in fslot); col_f / col_b hold the ranges of the others
  const int2* bslot = nullptr;
  // triangular solves with an LDS window of `window` rows: the entries of L (by row) and
  // of U (by row) that reach farther are gathered when their row enters the window
  const int* fwd_ptr = nullptr;  // n + 1
  const int* fwd_pos = nullptr;  // packed position of L(r, j)
  const int* fwd_col = nullptr;  // j
  const int* bwd_ptr = nullptr;  // n + 1
  const int* bwd_pos = nullptr;  // packed position of U(r, j)
  const int* bwd_col = nullptr;  // j
  int window = 0;                // window W the far lists were built for (ring rows W + kRingGap + 1)
  // narrow tiles: back substitution by levels (null: column by column).  Steps of kBackItems
  // items per lane group, item i of step q for group g at back_items[(q * kBackItems + i) * G + g];
  // back_steps is a multiple of 2 kBackAhead; see kBack* for the encoding
  const int2* back_items = nullptr;
  int back_steps = 0;
  // narrow tiles, refactor by parts (null: column by column with the fused forward substitution):
  // part_a[p] = {first entry in pent_pos / pent_div, entries, first chunk of part_slots, rounds};
  // pent_div = -1 for a U entry, -2 for a pivot, else the part-buffer offset of the entry's pivot;
  // part_x / part_w: the round items (x, word) of each chunk and lane group g at
  // (chunk * PG + g) * part_chunk_rounds(PG) + round, PG = part_groups (the solves' G, or 2 G for
  // quad tiles, whose refactor runs lane groups of two lanes); the forward substitution by levels
  // (fwd_items / fwd_steps, encoding as back_items with END keeping the sum)
  const int4* part_a = nullptr;
  const int* pent_pos = nullptr;
  const int* pent_div = nullptr;
  const int* part_x = nullptr;
  const int* part_w = nullptr;
  int n_parts = 0;
  int part_groups = 0;
  // 1 when the refactor stages structural-fill entries as zero itself (the fill kernel then leaves
  // those positions alone)
  int stage_zero = 0;
  const int2* fwd_items = nullptr;
  int fwd_steps = 0;
};

// Triangular-solve items {a, b} for a tile of TW scenarios (ring row stride S = TW + 1): a = kind << 29
// | payload, and
//   kBackEntry:   acc -= U(k, c) * x_c; payload = position of U(k, c); b = c * TW (x_c from y), or
//                 1 << 31 | slot * S (x_c from the LDS ring row `slot`)
//   kBackStart:   a piece of row k begins; payload = slot * S of the ring row its END writes; b = k * TW:
//                 acc = y_k (the row's last piece), or the zero row: acc = 0 (a partial piece)
//   kBackCombine: acc += the partial in ring row `slot`; payload 0; b = 1 << 31 | slot * S
//   kBackEnd:     the piece ends: a partial into its ring row; the row's last piece x_k = acc / U(k, k)
//                 (forward substitution: acc) into y and its ring row; payload = position of U(k, k)
//                 (0 for a partial and in the forward substitution); b = the zero row (at most one
//                 END per lane group and step)
//   kBackPad:     nothing; payload 0, b = the zero row
// The zero row is 1 << 31 | (R - 1) * S, the last of the R ring rows, which the kernel clears.
// The kernel loads a step's values at the end of the step kBackAhead before it.  Rows with more than
// kSolvePiece entries are cut into at most kSolvePieces pieces that run in different lane groups.
constexpr int kBackItems = 4;
constexpr int kBackAhead = 4;
constexpr int kBackEntry = 0, kBackStart = 1, kBackEnd = 2, kBackPad = 3, kBackCombine = 4;
constexpr int kSolvePiece = 8;
constexpr int kSolvePieces = 16;

// The linear initial guess system of a topology (complex, m = number of PQ buses): recipes
// for its packed L+U values and right-hand side.  Its LU pattern lives in a second DevTopo.
struct DevLg {
  int m = 0;
  int nnz_lu = 0;
  const int* src = nullptr;      // nnz_lu: Y-bus entry index of the value, or -1 (fill)
  const int* dbus = nullptr;     // nnz_lu: bus whose load admittance joins this diagonal, or -1
  const int* row_bus = nullptr;  // m: bus of unknown r
  const int* row_piv = nullptr;  // m: pivot row of equation r
  const int* row_x = nullptr;    // m: pivot column holding unknown r
  const int* rhs_ptr = nullptr;  // m + 1
  const int* rhs_e = nullptr;    // Y-bus entries coupling r to a bus of known voltage
  const int* rhs_j = nullptr;    // that bus
};

// Per-batch device workspace (sized for n_tiles tiles).
struct DevWork {
  int tile = 16;  // scenarios per tile (4, 8 or 16)
  int n_scen = 0;
  int n_tiles = 0;
  int max_iter = 0;
  double tol = 0.0;
  // bus arrays, interleaved [tile][n_bus][TW]
  double* vm = nullptr;
  double* va = nullptr;
  double* vr = nullptr;
  double* vi = nullptr;
  double* ps = nullptr;
  double* qs = nullptr;
  double* pc = nullptr;
  double* qc = nullptr;
  // unknown arrays, interleaved [tile][n][TW]
  double* f = nullptr;    // FP64 mismatch in pivot-row order
  double* dx = nullptr;   // FP64 Newton step in pivot-column order (mixed mode)
  void* y = nullptr;      // LU-precision right-hand side / solution
  void* lu = nullptr;     // packed L+U values, [tile][nnz_lu][TW]
  // per scenario
  int* status = nullptr;
  int* iters = nullptr;
  unsigned long long* maxmis = nullptr;  // max |mismatch| as ordered bits of a non-negative double
  double* history = nullptr;             // [n_scen][max_iter]
  int* running = nullptr;                // [max_iter] scenarios still running after each decision
  int stall_check = 0;                   // mixed precision: stop a scenario whose mismatch stops falling
  double* vm0 = nullptr;                 // starting point kept for the FP64 re-solve (mixed mode)
  double* va0 = nullptr;
  int* fallback = nullptr;               // [1] scenarios handed to the FP64 re-solve
  const int* outage = nullptr;           // [n_scen] branch out of service per scenario, -1 none (or null)
  int* iter_dev = nullptr;               // [1] the Newton iteration, when kept on the device (else an argument)
  // mixed precision: per tile, 1 when the tile runs this Newton step's refinement sweep (written by
  // k_dv before the correction solve, read by it and by k_update; a status change made by the
  // correction solve does not alter it)
  int* swept = nullptr;
  // per scenario (or null): 1 when the scenario's first Jacobian could not be factored (the
  // topology's pivot order came from a later scenario); it is Singular after its first mismatch
  const int* sing0 = nullptr;
};

// Launchers (gx_kernels.hip).  All run on `stream`; none synchronizes.  `fp32` selects the
// LU precision (float for the mixed-precision mode, double otherwise).
void launch_to_tiles(const double* src, double* dst, int n_scen, int n_tiles, int count, int tile,
                     hipStream_t stream);
void launch_from_tiles(const double* src, double* dst, int n_scen, int count, int tile, hipStream_t stream);
void launch_init(const DevTopo& t, const DevWork& w, hipStream_t stream);
// fill: also write the Jacobian values into the L+U storage (when t.y_lupos exists; else only the
// mismatch, and launch_fill writes them)
void launch_mismatch(const DevTopo& t, const DevWork& w, bool fp32, hipStream_t stream, bool fill = false);
void launch_decide(const DevTopo& t, const DevWork& w, int iter, hipStream_t stream);
void launch_fill(const DevTopo& t, const DevWork& w, bool fp32, hipStream_t stream);
void launch_lu(const DevTopo& t, const DevWork& w, bool fp32, bool factor, int iter, hipStream_t stream);
// Mixed precision: the FP64 residual of the refining scenarios' step as the right-hand side of the
// correction solve.
void launch_refine(const DevTopo& t, const DevWork& w, double* dvr, double* dvi, int iter, hipStream_t stream);
// (iter: the Newton step, for launches without a device iteration counter)
void launch_update(const DevTopo& t, const DevWork& w, bool fp32, bool mixed, int iter, hipStream_t stream);
void launch_finish(const DevWork& w, hipStream_t stream);
// Resident warm start of a topology: warm = |V| then theta of its buses (2 n_bus doubles), ok = 1 once
// it holds a solution.  launch_warm_start starts every scenario of w from it (when ok);
// launch_keep_warm stores w's first converged scenario into it (when there is one).
void launch_warm_start(const DevTopo& t, const DevWork& w, const double* warm, const int* ok, hipStream_t stream);
void launch_keep_warm(const DevTopo& t, const DevWork& w, double* warm, int* ok, hipStream_t stream);
// Mixed precision: restarts every scenario that did not converge from its starting point, for an
// FP64 re-solve (status back to running, count in w.fallback[0]).
void launch_fallback(const DevTopo& t, const DevWork& w, hipStream_t stream);
void launch_lane0(const double* src, double* dst, int count, bool to_tile, int stride, hipStream_t stream);
void launch_island_max(const DevTopo& t, const DevWork& w, const int* comp, int n_comp, unsigned long long* out,
                       hipStream_t stream);
void launch_linear_guess(const DevTopo& t, const DevLg& lg, const DevTopo& lgt, const DevWork& w, void* lu,
                         void* y, int* lg_status, int* lg_iters, hipStream_t stream);
// Dynamic LDS bytes of the factoring LU kernel for a topology (its window) and tile width:
// complex (linear guess), FP32 or FP64 values.
size_t lg_shmem(const DevTopo& lgt, int tile);
// One persistent LU launch over (request, tile) items of several requests (topos/works/items are
// device arrays; counter a device int the launch zeroes first); shmem = the largest request's.
void launch_lu_queue(int tile, const DevTopo* topos, const DevWork* works, const int2* items, int n_items,
                     int* counter, size_t shmem, int grid, bool fp32, bool factor, hipStream_t stream);
size_t lu_shmem_tile(const DevTopo& t, bool fp32, bool factor, int tile);
void launch_iter_advance(int* iter_dev, hipStream_t stream);
void launch_zero_ints(int* p, int n, hipStream_t stream);
size_t lu_shmem(const DevTopo& t, bool fp32, int tile);
// Allocation-free system solve: the static device array's address on the current device (looked up
// once per device), and the kernels that fill it from a system's values and read the solution out.
hipError_t scratch_base(void** out);
void launch_sys_stage(const int* pk_ptr, const int* pk_idx, const double* vals, int nlu, const int* prow,
                      const double* rhs, int n, double* lu, double* y, int* status, int* iters, hipStream_t stream);
void launch_sys_unpack(const double* y, const int* q, int n, double* x, hipStream_t stream);

}  // namespace gx
