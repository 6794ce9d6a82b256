/* gridox_hip — batched Newton-Raphson AC power flow on one MI355X (gfx950).
 *
 * C ABI: plain pointers and sizes only.  Every entry point states the reference interface it
 * replaces (file:line in m-mirz/gridoxide).  Host arrays are caller-owned; device memory is owned
 * by the library (per topology handle) except where an argument says "device pointer".
 *
 * Bus types use the reference's BusType (src/types.rs:5-9): 0 = Slack, 1 = PV, 2 = PQ.
 * Solve status uses SolveStatus (src/solver.rs:161-171): 0 = Converged, 1 = MaxIterationsReached,
 * 2 = Singular.
 *
 * Handles are not thread-safe: one handle per calling thread, like the reference's per-worker
 * PersistentSolver (src/batch.rs:233-246).
 */
#ifndef GRIDOX_HIP_H
#define GRIDOX_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ---- */
#define GX_OK 0
#define GX_ERR_ARG (-1)          /* invalid argument (sizes, indices, null pointers) */
#define GX_ERR_SINGULAR (-2)     /* structurally or numerically singular at symbolic/first factor */
#define GX_ERR_DEVICE (-3)       /* HIP runtime error (allocation, launch) */
#define GX_ERR_STATE (-4)        /* call out of order (e.g. pattern query before a factorization) */
#define GX_ERR_CAPACITY (-5)     /* caller buffer too small; required size returned where stated */
#define GX_ERR_UNSUPPORTED (-6)  /* topology exceeds a documented limit of the device path */
#define GX_ERR_BUDGET (-7)       /* a request does not fit the engine's device-memory budget now */

/* ---- solve status (SolveStatus, src/solver.rs:161-171) ---- */
#define GX_CONVERGED 0
#define GX_MAX_ITER 1
#define GX_SINGULAR 2

/* ---- numeric modes ---- */
#define GX_MODE_FP64 0   /* Jacobian, LU and solves in FP64 (the reference's arithmetic) */
#define GX_MODE_MIXED 1  /* FP64 mismatch/update/convergence; FP32 Jacobian/LU/solves; a scenario whose
                            mismatch is below 1e-2 pu refines its step with one sweep against the FP64
                            residual (an FP32 correction solve) */

/* ---- flags for gx_batch_solve ---- */
#define GX_DEVICE_PTRS 1 /* every array argument is a device pointer (inputs already in HBM) */
#define GX_INIT_LINEAR 2 /* start each scenario from linear_initial_guess (src/network.rs:627-678),
                            computed on the device from vm0/va0, as BatchSolver does
                            (src/batch.rs:268); without it vm0/va0 are used as given */
#define GX_INIT_WARM 4   /* start every scenario from the topology's resident warm start (the solution
                            a GX_KEEP_WARM solve left on the device) when it holds one, else from
                            vm0/va0 as given; not combined with GX_INIT_LINEAR.  The reference keeps a
                            topology's voltages between PersistentSolver solves (src/solver.rs:395-475) */
#define GX_KEEP_WARM 8   /* after the solve (a batch, or an engine request when it finishes) the
                            topology's resident warm start becomes the solution of its first
                            converged scenario, on the device (unchanged when none converged) */

const char* gx_version(void);
const char* gx_last_error(void);

/* Y-bus build.  Replaces build_ybus (src/network.rs:225-247), stamp_transformers
 * (src/network.rs:364-373) with branch_calc_param's closed (1,1) case (src/network.rs:339-360),
 * and YBus::finish (src/network.rs:52-71): pi-model lines (series 1/(r+jx), half the shunt
 * g + jb at each end; a line with from == to is a pure shunt), then transformers
 * (yff = (ys + ysh/2)/|tap|^2, yft = -ys/conj(tap), ytf = -ys/tap, ytt = ys + ysh/2); duplicate
 * entries summed in insertion order; CSR rows sorted by column.  Complex transformer arrays are
 * interleaved (re, im).  Returns nnz, or -(nnz needed) when cap is too small. */
int gx_ybus_build(int n_bus, int n_line, const int* line_from, const int* line_to, const double* r,
                  const double* x, const double* b_shunt, const double* g_shunt, int n_tr, const int* tr_from,
                  const int* tr_to, const double* y_series, const double* y_shunt, const double* tap, int cap,
                  int* rowptr, int* col, double* y_re, double* y_im);

/* ---- topology: Jacobian pattern + symbolic factorization, device-resident ---- */
typedef struct gx_topo gx_topo;

/* Replaces JacobianPattern::analyze (src/jacobian.rs:96-171) plus the symbolic half of
 * KluNativeSystem::new (build_csc_structure + analyze, src/klu_native/mod.rs:88-114,159-169;
 * analyze.rs:66-149) for one topology and bus-type assignment.  Host-only: device buffers are
 * allocated at the first solve.  Returns NULL on error (*err set). */
gx_topo* gx_topo_create(int n_bus, const int* bus_type, const int* rowptr, const int* col, const double* y_re,
                        const double* y_im, int* err);
void gx_topo_destroy(gx_topo* t);

/* sizes[0..7] = n_bus, nnz(Y), n unknowns, nnz(J), nnz(L+U) (0 before the first factor),
 * BTF blocks, factored (0/1), sum over U entries of |L(:,j)| (the refactor's multiply-adds). */
int gx_topo_sizes(const gx_topo* t, long long* sizes);

/* Jacobian pattern in the reference's emission order (H/N rows, then M/L rows;
 * src/jacobian.rs:128-168).  rows/cols sized nnz(J). */
int gx_topo_jacobian_pattern(const gx_topo* t, int* rows, int* cols);

/* The first numeric factorization that fixes the pivot order, done at one operating point:
 * KluNativeSystem::new's factor (src/klu_native/factor.rs:47-131, kernel.rs:187-262: partial
 * pivoting with diagonal preference, tol 1e-3, unscaled), evaluated on the Jacobian at
 * (vm, va, p_spec, q_spec).  newton_raphson_cached does this on the first iteration that needs a
 * solve (src/solver.rs:711-713); gx_batch_solve calls it implicitly with the first scenario's
 * starting point when the topology has not been factored yet. */
int gx_topo_factor_at(gx_topo* t, const double* vm, const double* va, const double* p_spec, const double* q_spec);

/* Accepts an externally computed pivot order instead (the route for a host that keeps its own
 * symbolic phase, e.g. KluNative's Symbolic.q and Numeric.pnum, src/klu_native/factor.rs:32-40):
 * column order q[k] and final row order pnum[k] of the Jacobian in the reference's row/column
 * numbering.  The L+U pattern follows from them. */
int gx_topo_set_pivot_order(gx_topo* t, const int* q, const int* pnum);

/* Orders after factoring: p/q (symbolic BTF+AMD row/column orders), r (block starts, nblocks+1
 * entries, buffer sized n+1), pnum (final pivot rows).  Any pointer may be NULL. */
int gx_topo_ordering(const gx_topo* t, int* p, int* q, int* r, int* pnum);

/* Packed L+U pattern in pivot space: cp[n+1] column starts, ulen[n] strictly-upper counts,
 * row[nnz(L+U)] row of each entry (column k: U rows ascending, diagonal, L rows ascending). */
int gx_topo_lu_pattern(const gx_topo* t, int* cp, int* ulen, int* row);

/* The level schedule of the back substitution that narrow tiles of `tile` scenarios use with a
 * solve ring of ring_rows rows (no counterpart in the reference; exposed so that tests can run it
 * on the CPU): 2 ints per item, items of step q, slot i, lane group g at index (q * 4 + i) * G + g
 * (G = 64 / tile); *steps = the number of steps.  Returns the ints written, or -(ints needed) when
 * cap is too small or items is NULL. */
int gx_topo_back_levels(const gx_topo* t, int tile, int ring_rows, int* items, long long cap, int* steps);

/* The forward substitution's level schedule for the same tiles (no counterpart in the reference;
 * for tests), in the same format: an END stores its sum (no division, payload 0). */
int gx_topo_forward_levels(const gx_topo* t, int tile, int ring_rows, int* items, long long cap, int* steps);

/* The refactor-by-parts schedule of narrow tiles of `tile` scenarios (no counterpart in the
 * reference; for tests).  sizes[4] receives the part count, the part entries, the ints of the round
 * slots and the staging capacity; when all four arrays are given (sized from a first call) they
 * receive part_a[4 * parts] (first entry, entries, first chunk, rounds of each part), pent_pos and
 * pent_div per entry (the packed L+U position; -1 U entry, -2 pivot, else the staging offset of the
 * column's pivot) and the slots ({x, word} per round and lane group, 16 rounds per chunk). */
int gx_topo_refactor_parts(const gx_topo* t, int tile, long long* sizes, int* part_a, int* pent_pos, int* pent_div,
                           int* slots);

/* ---- batched Newton-Raphson ---- */
/* Replaces BatchSolver::solve (src/batch.rs:198-289) driving PersistentSolver::solve_with_stats
 * (src/solver.rs:444-475) / newton_raphson_cached (src/solver.rs:639-736) once per scenario.
 * Inputs are scenario-major [n_scen][n_bus]: p_spec, q_spec (overrides already applied), and the
 * starting point vm0/va0 (used as given, or refined by GX_INIT_LINEAR).  Outputs:
 * vm, va [n_scen][n_bus]; status, iters [n_scen]; history [n_scen][max_iter] (max |mismatch| at
 * the top of each iteration; entries past iters are unspecified).  Non-convergence is a status,
 * not an error (src/batch.rs:193-197).  mode: GX_MODE_FP64 / GX_MODE_MIXED.  flags:
 * GX_DEVICE_PTRS.  stream: a hipStream_t (NULL = the null stream); the call returns after the
 * results are complete on that stream. */
int gx_batch_solve(gx_topo* t, int n_scen, const double* p_spec, const double* q_spec, const double* vm0,
                   const double* va0, double tol, int max_iter, int mode, int flags, double* vm, double* va,
                   int* status, int* iters, double* history, void* stream);

/* ---- islands ---- */
/* Island verdicts (IslandStatus, src/solver.rs:179-212). */
#define GX_ISLAND_SOLVABLE 0      /* exactly one Slack bus: solved; per-scenario status below */
#define GX_ISLAND_NO_REFERENCE 1  /* no Slack bus: every member pinned to V = 0, P = Q = 0 */
#define GX_ISLAND_AMBIGUOUS 2     /* two or more Slack buses: solved as is, reported as such */
/* Connected components of the Y-bus graph (connected_components + classify,
 * src/network.rs:121-193): comp[n_bus] = component of each bus (components numbered by their
 * lowest bus), verdict[n_comp] = GX_ISLAND_*.  Returns the number of components.  Members of an
 * unreferenced island become Slack buses at V = 0 with no injection in every solve
 * (mark_unreferenced_islands, src/network.rs:211-223). */
int gx_topo_islands(const gx_topo* t, int* comp, int* verdict);
/* Per-scenario, per-island outcome of the last gx_batch_solve on t (finish_island_reports +
 * resolve_pending, src/solver.rs:234-312): out[s][c] is GX_CONVERGED / GX_MAX_ITER / GX_SINGULAR
 * for a solvable island (its own max mismatch at the final voltages against tol, else the
 * scenario's status), or GX_ISLAND_STATUS_NO_REFERENCE / GX_ISLAND_STATUS_AMBIGUOUS.  out is a
 * host array [n_scen][n_comp]. */
#define GX_ISLAND_STATUS_NO_REFERENCE 3
#define GX_ISLAND_STATUS_AMBIGUOUS 4
int gx_batch_island_status(gx_topo* t, int* out);

/* ---- per-kernel timing (no counterpart in the reference: measurement support) ---- */
#define GX_K_MISMATCH 0 /* injections + mismatch */
#define GX_K_FILL 1     /* Jacobian values into the L+U storage */
#define GX_K_LU 2       /* numeric refactor + forward/back substitution */
#define GX_K_REFINE 3   /* mixed precision: FP64 residual of the step */
#define GX_K_SOLVE 4    /* mixed precision: correction solve with the FP32 factors */
#define GX_K_UPDATE 5   /* voltage update */
#define GX_K_COUNT 6
/* Reads and clears the accumulated HIP-event timings of t's launches since the last call
 * (ms[GX_K_COUNT] milliseconds summed per kernel class, launches[GX_K_COUNT]; either may be NULL),
 * then turns timing on or off for later gx_batch_solve calls.  Timing brackets each launch with
 * events on the solve's own stream. */
int gx_topo_profile(gx_topo* t, int enable, double* ms, int* launches);

/* Launch shape of later gx_batch_solve calls on t (no counterpart in the reference, whose CPU
 * workers have no such knob; this is the setting a tuner sweeps).  tile: scenarios per
 * wavefront tile, 2, 4, 8, 16 (0 = automatic: 8 from 8,192 scenarios, 4,096 in mixed precision,
 * else 2 while the tiles fit one wavefront per SIMD, else the narrowest tile whose wavefronts can
 * all be resident at once), or 64 (one lane per scenario).  8-scenario tiles run two scenarios per
 * lane (pair tiles).
 * ring_rows: rows of the LDS window that holds the right-hand side during the triangular solves,
 * a power of two in [64, 4096] (0 = the default for the tile); it is halved while the kernel's
 * LDS would exceed 160 KB.  Results do not depend on either setting beyond rounding of the
 * triangular-solve sums.  GX_ERR_ARG for other values. */
int gx_topo_set_launch(gx_topo* t, int tile, int ring_rows);

/* Lane groups of the refactor of mixed-precision 8-scenario tiles in later gx_batch_solve calls on
 * t (no counterpart in the reference).  0 (automatic) and 32: quad tiles, whose refactor runs 32
 * lane groups of two lanes with four FP32 scenarios per lane (the substitutions keep 16 groups of
 * four lanes, two scenarios per lane), when the topology's part schedule exists for 32 groups; 16:
 * the refactor on the substitutions' 16 groups.  Results are the same bits either way (each entry
 * takes its updates in the same order).  GX_ERR_ARG for other values. */
int gx_topo_set_refactor_groups(gx_topo* t, int groups);

/* Voltage-dependent ZIP injection terms of t's buses (Bus::zip_terms, src/types.rs:14-24, 40-41),
 * summed per bus and kind: zip[6*i .. 6*i+5] = (P, Q) of bus i's constant-power, constant-current
 * and constant-impedance terms in per unit at |V| = 1.  Every later solve evaluates
 * effective_injection (src/network.rs:600-611) with them: P_eff = p_spec + P_p + P_i*vm + P_z*vm^2
 * (Q alike), in the mismatch, the island check and the linear initial guess; the Jacobian keeps the
 * reference's constant-power form.  zip == NULL removes them.  Takes effect at the next solve
 * or engine submission; not to be called while an engine request on t is in flight. */
int gx_topo_set_zip(gx_topo* t, const double* zip);

/* Branches of t for per-scenario outages, in the argument layout of gx_ybus_build (lines, then
 * transformers; branch e < n_line is line e, else transformer e - n_line).  The reference's
 * Scenario::branch_outages (src/batch.rs:70-83) names branches but BatchSolver rejects them
 * (OutagesUnsupported, src/batch.rs:207-209); this is the N-1 extension that keeps the Y-bus
 * pattern: an outaged branch's four stamps (build_ybus, src/network.rs:225-247;
 * branch_calc_param, src/network.rs:339-360) are subtracted from that scenario's Y values, so the
 * symbolic factorization stays shared.  GX_ERR_ARG if a branch is not in t's Y-bus pattern. */
int gx_topo_set_branches(gx_topo* t, int n_line, const int* line_from, const int* line_to, const double* r,
                         const double* x, const double* b_shunt, const double* g_shunt, int n_tr,
                         const int* tr_from, const int* tr_to, const double* y_series, const double* y_shunt,
                         const double* tap);

/* 1 if removing branch `branch` would split its island (a bridge; such outages are rejected), else
 * 0; negative error code. */
int gx_topo_branch_is_bridge(const gx_topo* t, int branch);

/* gx_batch_solve with one branch out of service per scenario: outage[s] is a branch index of
 * gx_topo_set_branches or -1 (length n_scen; a device pointer with GX_DEVICE_PTRS).  GX_ERR_ARG for
 * an index out of range or a bridge, GX_ERR_STATE if branches were never set.  outage == NULL is
 * gx_batch_solve. */
int gx_batch_solve_outages(gx_topo* t, int n_scen, const double* p_spec, const double* q_spec, const double* vm0,
                           const double* va0, const int* outage, double tol, int max_iter, int mode, int flags,
                           double* vm, double* va, int* status, int* iters, double* history, void* stream);

/* Number of scenarios of the last GX_MODE_MIXED gx_batch_solve on t that were solved again from
 * their starting point with the FP64 path: those whose mismatch stopped falling once below
 * 1e-4 pu (the FP32 factors limit the step) and those with a singular FP32 factor.  Their
 * reported status, iterations and history are the FP64 solve's.  A mixed-precision run that
 * neither converges nor stalls keeps its MaxIterationsReached result. */
int gx_batch_fallbacks(const gx_topo* t);

/* ---- LinearSolver-level system ---- */
typedef struct gx_sys gx_sys;
/* Replaces LinearSolver::new for KluNative (src/solver.rs:35-57, src/klu_native/mod.rs:159-169):
 * triplets (duplicates summed), analyze + first factorization with pivoting.  NULL when
 * structurally or numerically singular (the reference returns None). */
gx_sys* gx_sys_create(int n, int nnz, const int* rows, const int* cols, const double* vals);
/* Replaces factor_and_solve_values (src/klu_native/mod.rs:185-210): values positionally matching
 * the creation triplets; refactor on the device with the pivot order kept, then solve.
 * Returns GX_OK, or GX_SINGULAR when a pivot is zero or the solution is not finite (None). */
int gx_sys_factor_and_solve_values(gx_sys* s, const double* vals, const double* rhs, double* x);
void gx_sys_destroy(gx_sys* s);
/* No counterpart in the reference: for callers that cannot allocate device memory while they solve
 * (code under HIP graph capture, tight serving loops).  gx_sys_factor_and_solve_values with device
 * pointers -- vals (the creation triplets' values, positionally), rhs and x (length n) -- on `stream`
 * (a hipStream_t, NULL = the null stream), returning after x is written: GX_OK, or GX_SINGULAR (x is
 * then unspecified).  It allocates nothing: the system's read-only tables are uploaded by
 * gx_sys_create (when a GPU is present; else GX_ERR_STATE), and the L+U values, right-hand side and
 * status live in one static 64 MiB device array per GPU, so a system may hold at most about
 * 2,000,000 L+U entries (32 B each plus 32 B per row; GX_ERR_CAPACITY beyond).  Every call on a GPU
 * shares that array: calls that overlap on one GPU (other threads or streams) race. */
int gx_sys_factor_and_solve_values_noalloc(gx_sys* s, const double* vals, const double* rhs, double* x,
                                           void* stream);
/* Symbolic/pivot orders of a system (the Symbolic p/q/r and Numeric pnum of
 * src/klu_native/analyze.rs:22-54 and factor.rs:32-40).  p, q, pnum sized n; r sized n+1;
 * *nblocks receives the BTF block count.  Returns nnz(L+U) (strictly-triangular parts plus the
 * diagonal, the whole matrix factored as one block in the given order) or an error code. */
int gx_sys_ordering(const gx_sys* s, int* p, int* q, int* r, int* pnum, int* nblocks);

/* ---- request engine (uneven batches and serving) ----
 * No counterpart in the reference, whose BatchSolver runs one batch over one topology on rayon
 * workers (src/batch.rs:198-289).  Many requests -- each one topology + its scenarios, as in
 * gx_batch_solve (a topology used by an engine keeps the engine's tile width) -- are in flight
 * at once: every step runs one Newton iteration of every active request, with ONE persistent LU
 * launch over the tiles of all of them (a device work queue, heaviest first).  Requests are
 * admitted between steps and leave when finished; results per request equal gx_batch_solve's
 * for the same inputs, tile width and pivot order. */
typedef struct gx_engine gx_engine;
/* tile: 2, 4, 8 (pair tiles, as gx_batch_solve's), 16 or 64 (0: 4).  hbm_budget: bytes of request workspaces the engine may hold (0:
 * unlimited).  stream: HIP stream of all its work (NULL: its own).  NULL on error (*err set). */
gx_engine* gx_engine_create(int tile, int flags, long long hbm_budget, void* stream, int* err);
void gx_engine_destroy(gx_engine* e);
/* Device bytes a request would hold (an estimate before its topology's first factorization). */
long long gx_engine_request_bytes(const gx_topo* t, int n_scen, int max_iter, int mode, int flags, int tile);
/* Device bytes held by the engine's requests (active and finished-but-not-collected). */
long long gx_engine_bytes(const gx_engine* e);
/* Adds a request (arguments as gx_batch_solve_outages; flags GX_DEVICE_PTRS, GX_INIT_LINEAR).  Returns
 * its id (>= 0), GX_ERR_BUDGET if it does not fit the budget now (retry after collecting), or an
 * error code. */
int gx_engine_submit(gx_engine* e, gx_topo* t, int n_scen, const double* p_spec, const double* q_spec,
                     const double* vm0, const double* va0, const int* outage, double tol, int max_iter, int mode,
                     int flags);
/* One Newton step of every active request; returns how many remain active. */
int gx_engine_step(gx_engine* e);
/* Up to max_steps steps with each step captured once in a hipGraph and replayed (re-captured when
 * the set of requests changes); returns how many remain active. */
int gx_engine_run(gx_engine* e, int max_steps);
/* 1 when request id has finished: its results (as gx_batch_solve's outputs, host or device
 * pointers as submitted) are written and the request is released; 0 while it runs. */
int gx_engine_collect(gx_engine* e, int id, double* vm, double* va, int* status, int* iters, double* history);
/* Scenarios of a finished mixed-precision request that were re-solved with the FP64 path. */
int gx_engine_fallbacks(const gx_engine* e, int id);

#ifdef __cplusplus
}
#endif
#endif
