/* librumgpu: the MI355X (gfx950) BM-ADMM and Flow-CG engines of the RUM distance, as a C ABI.
 *
 * Plain pointers and sizes only. Every entry returns an int status:
 *   0 ok, 1 invalid argument, 2 out of device memory, 3 HIP error, 4 unsupported size;
 * rumgpu_last_error() returns the message of the last failure in this process.
 * Host arrays are read or written synchronously before the call returns; device memory stays
 * behind the opaque handles. One context = one device and one HIP stream; a handle is not
 * thread-safe. Row arrays: menu = bitmask of the offered set (bit a = alternative a), alt = the
 * chosen alternative (0-based), one entry per observed non-singleton (menu, alternative) cell.
 *
 * Reference interfaces each entry replaces are cited as path:line into the reference
 * repository (mingxiz/rum-testing-gpu).
 */
#ifndef RUMGPU_H
#define RUMGPU_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rumgpu_ctx rumgpu_ctx;
typedef struct rumgpu_bm rumgpu_bm;
typedef struct rumgpu_cg rumgpu_cg;
typedef struct rumgpu_mlp rumgpu_mlp;

/* ---- context, device, errors (replaces CuPy's device/pool handling and machine.py:25-61,
 *      machine.py:156 gpu_free_bytes) ---- */
int rumgpu_ctx_create(int device, rumgpu_ctx** out);
int rumgpu_ctx_destroy(rumgpu_ctx* ctx);
int rumgpu_mem_info(int device, size_t* free_bytes, size_t* total_bytes);
int rumgpu_device_props(int device, char* name, int name_len, char* arch, int arch_len, size_t* total_mem,
                        int* cu_count);
int rumgpu_device_count(int* count);
int rumgpu_last_error(char* buf, int len);
int rumgpu_version(void);

/* ---- BM-ADMM (replaces solve_kadmm_v2, solver/bm_admm/solver.py:148-377, its kernels
 *      :33-139, bm_admm/kernels.py:23-35 and bm_admm/butterfly2.py:206-247) ----
 * dtype_bytes 8 = float64 iterates, 4 = float32 iterates. (e0, e1, Q row-major) = eigh of
 * [[1,-1],[-1,1+p^2]] and gamma = beta2/beta1 (solver.py:206-214). flags bit 0 = the plain
 * reference schedule (separate elementwise kernels, four full butterflies per iteration)
 * instead of the fused window passes. The object starts cold: x2 = prox(xhat), s = y = z = 0
 * (solver.py:264-267). */
int rumgpu_bm_create(rumgpu_ctx* ctx, int n, const uint64_t* menu, const int32_t* alt, const double* p, int64_t M,
                     int dtype_bytes, double beta, double alpha, double precond_p, double e0, double e1, double gamma,
                     const double* Q, int flags, rumgpu_bm** out);
int rumgpu_bm_destroy(rumgpu_bm* bm);
/* new probabilities on the same rows (a bootstrap draw): Session.solve(p), rum_solver.py:598 */
int rumgpu_bm_set_p(rumgpu_bm* bm, const double* p);
/* cold start: solver.py:264-267 */
int rumgpu_bm_init_cold(rumgpu_bm* bm);
/* state (x2, s, y, z) as float64 host slabs of n 2^(n-1) values: init_state, solver.py:256-262 */
int rumgpu_bm_set_state(rumgpu_bm* bm, const double* x2, const double* s, const double* y, const double* z);
int rumgpu_bm_get_state(rumgpu_bm* bm, double* x2, double* s, double* y, double* z);
/* device-side state copy between two objects on the same rows (dtype conversion allowed): the
 * float32 -> float64 handover of solve_kadmm_hybrid, solver.py:380-425 */
int rumgpu_bm_copy_state(rumgpu_bm* dst, rumgpu_bm* src);
/* keep the current state on the device / restart from it: Session keep= / warm=,
 * rum_solver.py:922-924 (_run_bm_admm keep) and :914-925 (_bm_session warm) */
int rumgpu_bm_keep(rumgpu_bm* bm);
int rumgpu_bm_load_kept(rumgpu_bm* bm);
/* run iterations (one_iter, solver.py:240-254); the state afterwards may hold a deferred s, y
 * update, settled by rumgpu_bm_finalize or any state read */
int rumgpu_bm_run(rumgpu_bm* bm, int64_t iters);
/* capture `len` steady-state iterations as one HIP graph, replayed by rumgpu_bm_run when the
 * count is a multiple of len (solver.py:269-292); len 0 drops it */
int rumgpu_bm_set_graph(rumgpu_bm* bm, int64_t len);
/* change the penalty (beta1 = beta2 = beta) between iterations; the unscaled multipliers carry
 * over. The reference fixes beta per solve (solver.py:148-154); this is the hook of an adaptive
 * penalty (residual balancing on the certified gap). */
int rumgpu_bm_set_beta(rumgpu_bm* bm, double beta);
/* window plan of the fused passes: lo_bits (1..13) at bit 0 (lo_rl register bits), the rest in
 * windows of <= 8 bits (hi_rl register bits); a plan whose tile needs more threads (512) or LDS
 * than a workgroup may have on the device is refused (status 1). The reference plans its butterfly passes by a calibrated
 * DP (common/butterfly.py:125-195); here the autotuner (solver/bm_admm/autotune.py) times plans
 * per n. Results are the same for every plan up to rounding order. */
int rumgpu_bm_set_plan(rumgpu_bm* bm, int lo_bits, int lo_rl, int hi_rl);
int rumgpu_bm_finalize(rumgpu_bm* bm);
/* phi = ||x2[obs] - p|| (solver.py:321), compensated float64 sum */
int rumgpu_bm_phi(rumgpu_bm* bm, double* phi);
/* float64 push to a feasible RUM point: ub and its fitted probabilities (input row order, may
 * be NULL). push_to_ub64, common/transforms.py:94-98 + arcflow_kernels.py:415-434 */
int rumgpu_bm_push_ub64(rumgpu_bm* bm, double* ub, double* p_fit);
/* the fitted probabilities of the last push */
int rumgpu_bm_get_fit(rumgpu_bm* bm, double* p_fit);
/* float64 Frank-Wolfe lower bound at x2: fw_lb_rows_gpu, common/transforms.py:101-114 */
int rumgpu_bm_lb64(rumgpu_bm* bm, double* lb);
/* x2 at the observed rows (input order), float64 */
int rumgpu_bm_x_obs(rumgpu_bm* bm, double* out);
int rumgpu_bm_info(rumgpu_bm* bm, int64_t* iters, int* nwin, int* win_bits);
/* device time per fused pass kind over `iters` iterations (HIP events between launches on the
 * context stream): ms_by_kind[7], launches_by_kind[7] for SY, RHS, D, XK, prox, middle-window
 * transform, one-window pass. No reference counterpart: the measurement behind the roofline. */
int rumgpu_bm_profile(rumgpu_bm* bm, int64_t iters, double* ms_by_kind, int64_t* launches_by_kind);

/* ---- Draw slots: many bootstrap draws per launch (replaces the per-draw Session.solve loop of
 * Bootstrap._solve, inference/bootstrap.py:308-335, warm from the kept p-hat state, :318).
 * A batch object holds `slots` independent ADMM states on the same rows; rumgpu_bm_run advances
 * every slot, each slot gets its own p, its own start and its own float64 certificates. The
 * single-solve entry points above act on slot 0; set_state, copy_state and load_kept refuse a
 * batch object (use slot_load). */
int rumgpu_bm_create_batch(rumgpu_ctx* ctx, int n, const uint64_t* menu, const int32_t* alt, const double* p,
                           int64_t M, int dtype_bytes, double beta, double alpha, double precond_p, double e0,
                           double e1, double gamma, const double* Q, int slots, rumgpu_bm** out);
/* p of one slot (input row order): Session.solve(P_b), bootstrap.py:318 */
int rumgpu_bm_slot_set_p(rumgpu_bm* bm, int slot, const double* p);
/* warm start of one slot from the state `src` kept (rumgpu_bm_keep): init_state=warm,
 * rum_solver.py:914-925; any dtypes; the slot's next iteration starts from that state */
int rumgpu_bm_slot_load(rumgpu_bm* bm, int slot, rumgpu_bm* src);
/* cold start of one slot: solver.py:264-267 */
int rumgpu_bm_slot_cold(rumgpu_bm* bm, int slot);
/* corrections added at the observed cells of one slot's x2 and z (either may be null), and the
 * observed cells of x2 and z read back (float64, input row order). No reference counterpart: the
 * learned warm start's output and its training targets. */
int rumgpu_bm_slot_add_obs(rumgpu_bm* bm, int slot, const double* dx2, const double* dz);
int rumgpu_bm_slot_get_obs(rumgpu_bm* bm, int slot, double* x2_obs, double* z_obs);
/* float64 certificates of every slot: phi[slots] (solver.py:321), ub[slots] (push_to_ub64,
 * transforms.py:94-98, fitted point kept per slot), lb[slots] when want_lb (fw_lb_rows_gpu,
 * transforms.py:101-114) */
int rumgpu_bm_certify(rumgpu_bm* bm, int want_lb, double* ub, double* lb, double* phi);
/* the fitted point of one slot's last certificate: p_fit, rum_solver.py:884-889 */
int rumgpu_bm_slot_fit(rumgpu_bm* bm, int slot, double* p_fit);
/* ---- Device draw queue of a batch object: the open draws' loop of Bootstrap._solve
 * (inference/bootstrap.py:308-335; each draw's stops as Session.solve applies them through
 * bm_admm/solver.py:294-377) run with retirement and refill on the device. A cycle is `check`
 * iterations of every slot, their float64 certificates (the lower bound every lb_every-th cycle),
 * the stop tests of every live slot in slot order (ub below `below`, lb at or above `above`,
 * ub - lb within abs_tol or rel_tol * lb, the host's time stop, max_iter) and the refill of each
 * retired slot with the next pending draw, warm from the state `src` kept (plus corrections at the
 * observed cells when corr is 1: dx2, 2: dz, 3: both). Cycles replay as HIP graphs; the host only
 * tops up the ring of pending draws and reads results between runs of many cycles. */
int rumgpu_bm_queue_start(rumgpu_bm* bm, rumgpu_bm* src, int64_t total, int64_t capacity, int check, int lb_every,
                          int64_t max_iter, double abs_tol, double rel_tol, int corr);
/* k pending draws: p (k x M, input row order), thresholds on the non-singleton part (NaN: none),
 * corrections (k x M) when the queue takes them; at most `capacity` untaken draws at a time */
int rumgpu_bm_queue_push(rumgpu_bm* bm, int64_t k, const double* p, const double* below, const double* above,
                         const double* dx2, const double* dz);
/* `cycles` cycles; status[5] = draws taken, pushed, retired, live slots, checks */
int rumgpu_bm_queue_run(rumgpu_bm* bm, int64_t cycles, int64_t* status);
/* per draw (sequence numbers first..first+count-1): lb, ub of the non-singleton part, iterations,
 * retirement rank, stop (0 open, 1 threshold, 2 tolerance, 3 time, 4 max_iter); outputs may be null */
int rumgpu_bm_queue_results(rumgpu_bm* bm, int64_t first, int64_t count, double* lb, double* ub, int64_t* iters,
                            int64_t* order, int32_t* stop);
/* per slot: the draw it holds (-1 idle) and its iterations */
int rumgpu_bm_queue_slots(rumgpu_bm* bm, int64_t* draw, int64_t* iters);
/* time stops (draw_time, rum_solver.py:914-925 max_time per draw): kill[slot] != 0 stops that
 * slot's draw at the next decision */
int rumgpu_bm_queue_expire(rumgpu_bm* bm, const int32_t* kill);
int rumgpu_bm_queue_end(rumgpu_bm* bm);
/* building blocks exposed for tests: one full (a,b,c,d) butterfly of a host slab
 * (common/butterfly.py:99-140 butterfly_fast), and the prox of a host slab (solver.py:66-139) */
int rumgpu_butterfly64(rumgpu_ctx* ctx, int n, double* v, double a, double b, double c, double d);
int rumgpu_butterfly32(rumgpu_ctx* ctx, int n, float* v, float a, float b, float c, float d);
int rumgpu_prox_slab(rumgpu_bm* bm, const double* xk, double* out);

/* float64 Frank-Wolfe lower bound at any point y on the rows, with its own scratch:
 * fw_lb_rows_gpu, common/transforms.py:101-114 (flow_cg's final bound, rum_solver.py:928-938) */
int rumgpu_lb64(rumgpu_ctx* ctx, int n, const uint64_t* menu, const int32_t* alt, int64_t M, const double* y,
                const double* p, double* lb);
/* The same lower bound with no allocation. No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). menu_dev, alt_dev, y_dev and p_dev are M rows in device memory of ctx's device;
 * lb_dev is one double in device memory. Every buffer rumgpu_lb64 allocates per call lives instead in
 * one fixed device array of the library, sized for n <= 18 (22,028,800 bytes, held by each device the
 * library runs on until the process ends); larger n returns status 4. Work runs on ctx's stream and
 * has finished when the call returns. Calls on one device share that array: calls that overlap on one
 * GPU (from several threads or contexts) race. A row whose alternative is outside 0..n-1 or not in its
 * menu, or whose menu has bits at n or above, makes lb = NaN. Same kernels and order as rumgpu_lb64. */
int rumgpu_lb64_noalloc(rumgpu_ctx* ctx, int n, const uint64_t* menu_dev, const int32_t* alt_dev, int64_t M,
                        const double* y_dev, const double* p_dev, double* lb_dev);

/* ---- Flow-CG pricing (replaces bf_update, flow_cg/bf_pricing.py:8-47 with the sweep and
 *      backtrace of flow_cg/solver.py:168-200, and StreamPricer, flow_cg/pricing_stream.py:267-412)
 * Rows are given x-major (stable by alternative, cg_core.py:184-195): offsets[x]..offsets[x+1]
 * are alternative x's rows, masks their menus.
 * layout 0 = mask (one float32 per subset, the reference's row-order sums: exact for any scores)
 * layout 1 = slab (scores scattered on the slab and summed over subsets by the butterfly, then
 *            the layered max: the same values as mask when every partial sum is exact)
 * layout 2 = stream (two live layers by colex rank), with dtype 0 f32, 1 u16, 2 u24 storage;
 * k_pred: first layer whose predecessors are stored (stream), <= 0 = n - 10. */
int rumgpu_cg_create(rumgpu_ctx* ctx, int n, const uint64_t* masks_x, const uint64_t* offsets, int64_t M, int layout,
                     int dtype, int k_pred, rumgpu_cg** out);
int rumgpu_cg_destroy(rumgpu_cg* cg);
/* one pricing sweep: scores (float32, x-major) -> vmax (float64 decode), rank_pos[n] (position
 * of each alternative, 0 = best), margin (u16/u24 quantization bound, 0 otherwise). scores_x =
 * NULL prices the scores rumgpu_cg_cols_scores left on the device. */
int rumgpu_cg_price(rumgpu_cg* cg, const float* scores_x, double* vmax, int32_t* rank_pos, double* margin);
/* D pricing sweeps in one pass over the lattice (mask layout): scores is D x M (draw-major, x-major
 * rows), vmax[D], rank_pos[D x n]. Draw d's result equals rumgpu_cg_price on row d of scores bit
 * for bit. No reference counterpart: the batched bootstrap prices the open draws' columns
 * together instead of one bf_update sweep per draw (flow_cg/solver.py:168-200). */
int rumgpu_cg_price_batch(rumgpu_cg* cg, int D, const float* scores, double* vmax, int32_t* rank_pos);
/* float64 max over rankings of sum_j r_j [the ranking picks row j's alternative] for r given per
 * row (x-major): the longest path behind the Frank-Wolfe lower bound of fw_lb_rows_gpu
 * (common/transforms.py:101-114), flow_cg's final bound (rum_solver.py:928-938), computed on the
 * pricer's lattice instead of a float64 slab: mask layout (2^n float64 scratch, kept) or stream
 * layout (two float64 colex layers, sweep_max64 / fw_lb_stream64, pricing_stream.py:190-248,
 * transforms.py:117-129) */
int rumgpu_cg_vmax64(rumgpu_cg* cg, const double* r_x, double* vmax);
/* the same float64 longest path for D residual vectors at once (mask layout): r is D x M (draw-major,
 * x-major rows), vmax[D]; draw d's value equals rumgpu_cg_vmax64 on row d bit for bit. No reference
 * counterpart: the batched bootstrap takes the final bounds of the draws that finish in one round
 * together instead of one fw_lb_rows_gpu per draw (transforms.py:101-114). */
int rumgpu_cg_vmax64_batch(rumgpu_cg* cg, int D, const double* r, double* vmax);
/* rumgpu_cg_price_batch for D draws and rumgpu_cg_vmax64_batch for L more, the same values, in one
 * call: on the one-alternative window-pass sweep (D <= 32, L <= 8) the float64 maxima of the draws
 * that finish come out of the same sweep that prices the others. No reference counterpart (the
 * reference prices each draw and takes each draw's final bound apart, bootstrap.py:308-335). */
int rumgpu_cg_price_batch_lb(rumgpu_cg* cg, int D, const float* scores, double* vmax, int32_t* rank_pos, int L,
                             const double* r, double* vmax64);
int rumgpu_cg_info(rumgpu_cg* cg, size_t* device_bytes, double* last_kernel_ms);
/* How the mask layout sweeps (rumgpu_cg_price, _price_batch, _vmax64): mode 0 (the default) =
 * layer by layer for any rows, in the reference's order (bf_update, flow_cg/bf_pricing.py:7-47);
 * mode 1 = when every row belongs to one alternative a (a panel of one alternative's choice
 * shares), the same V[full] and ranking from a max over W[T] = a's row-order float32 sums on the
 * 2^(n-1) subsets T of the other alternatives, with one bit per node and draw closed under
 * supersets for the path; 4 bytes per node instead of V and pred for 2^n nodes (refused above
 * n = 20); mode 2 = the same rule in window passes of other kernels (W with block maxima, the bits
 * closed in 12- and 8-bit windows, a block per draw for the path; n <= 26, at most 4,096 rows).
 * No reference counterpart: the reference's sweep does not special-case one alternative.
 * sweep_info reports the alternative mode 1 or 2 uses, or -1 when the layered sweep runs. */
int rumgpu_cg_set_sweep(rumgpu_cg* cg, int mode);
int rumgpu_cg_sweep_info(rumgpu_cg* cg, int* one_alternative);
/* Tile bits L of the slab layout's layered max (2..10, 0 = the built-in choice): the subsets are
 * worked in tiles of 2^L, one wavefront per tile, one launch per popcount of the other n - L bits.
 * Results are the same for every L. No reference counterpart: a tuning knob of this engine's
 * tiling of the layered DP (flow_cg/solver.py:168-200 sweeps one layer at a time). */
int rumgpu_cg_set_slab_tile(rumgpu_cg* cg, int bits);

/* Column store of one column-generation run, on the pricer's device and rows. Replaces the host
 * column bookkeeping of solve_cgbfqp_cholbitset (flow_cg/solver.py:92-115, :145, :240): the packed
 * column bitsets and Gram rows of BitsetGramMaintainer (flow_cg/qp_bitset_incremental.py:36-79),
 * p_out = V lam of PoutMaintainer (:99-130) and the ranking-to-column map (map_ranking_to_choice,
 * flow_cg/cg_core.py:64-80). Columns are bitsets over the x-major rows: bit r % 64 of word r / 64,
 * W = ceil(M / 64) words per column.
 * cols_init: p_x[M] (float64, x-major) and room for cap columns; drops stored columns.
 * cols_load: k columns (k x W words) replace the stored ones (a warm start's columns).
 * cols_add: the column of a ranking (rank_pos[n], a permutation; row r is 1 iff its alternative
 *   is its menu's best-ranked member) is appended; gram[K + 1] receives its inner products with
 *   columns 0..K (itself last, exact integers), h its inner product with p (a fixed-order sum
 *   within each block of 256 rows, then over the blocks in order), bits[W] the column unless null.
 * cols_scores: p_out = sum_t lam[t] col[idx[t]] (in t order) and scores = float32(p - p_out) stay
 *   on the device; rumgpu_cg_price(cg, NULL, ...) then prices those scores.
 * cols_pout: p_out[M] of the last cols_scores. */
int rumgpu_cg_cols_init(rumgpu_cg* cg, const double* p_x, int64_t cap);
int rumgpu_cg_cols_load(rumgpu_cg* cg, int64_t k, const uint64_t* bits);
int rumgpu_cg_cols_add(rumgpu_cg* cg, const int32_t* rank_pos, double* gram, double* h, uint64_t* bits);
int rumgpu_cg_cols_scores(rumgpu_cg* cg, int32_t nA, const int32_t* idx, const double* lam);
int rumgpu_cg_cols_pout(rumgpu_cg* cg, double* p_out);

/* Flow-CG restricted-master QP on the host: min 1/2 lam'Q lam - h'lam, lam >= 0, sum lam = 1, by
 * the active set of solve_qp_native (common/qp_native.py:37-100) with its warm working set
 * A_init, then the simplex projection (:20-33). Q is K x K row-major with leading dimension ldq.
 * Writes lam[K] and the support (weights > 1e-9) into support[K], its size into n_support. */
int rumgpu_qp_simplex(int K, const double* Q, int64_t ldq, const double* h, const int32_t* A_init, int nA,
                      int max_outer, double tol, double reg, double* lam, int32_t* support, int* n_support);
/* the same with a state that keeps the last working set and its Cholesky factor: a column
 * generation whose next call starts from that working set (Q only grew) skips the refactoring */
typedef struct rumgpu_qp rumgpu_qp;
int rumgpu_qp_create(rumgpu_qp** out);
int rumgpu_qp_destroy(rumgpu_qp* state);
int rumgpu_qp_simplex_warm(rumgpu_qp* state, int K, const double* Q, int64_t ldq, const double* h,
                           const int32_t* A_init, int nA, int max_outer, double tol, double reg, double* lam,
                           int32_t* support, int* n_support);
/* ---- The host side of one column-generation run between two sweeps (the loop body of
 * flow_cg/solver.py:92-240, with qp_bitset_incremental.py:36-130's bitset Gram rows): ranking
 * columns as bitsets over the x-major rows (cg_core.py:63-80: a row is 1 iff its alternative is
 * its menu's best-ranked member), the integer Gram matrix, h = V'p (rows ascending), the master
 * QP warm from the last working set and factor (rumgpu_qp_simplex_warm), p_out = V lam over lam's
 * support (columns ascending), and the next sweep's scores float32(p - p_out), one call per round.
 * Host code; no device memory. */
typedef struct rumgpu_cgr rumgpu_cgr;
int rumgpu_cgr_create(int n, int64_t M, const uint64_t* menu_x, const int32_t* alt_x, const double* p_x, int max_K,
                      rumgpu_cgr** out);
int rumgpu_cgr_destroy(rumgpu_cgr* c);
int rumgpu_cgr_add_ranking(rumgpu_cgr* c, const int32_t* rank_pos);
int rumgpu_cgr_add_column(rumgpu_cgr* c, const uint8_t* col);
/* dst's columns become src's (same rows), with h recomputed from dst's p */
int rumgpu_cgr_copy_columns(rumgpu_cgr* dst, const rumgpu_cgr* src);
int rumgpu_cgr_set_lambda(rumgpu_cgr* c, const double* lam, int K, const int32_t* working_set, int n_working);
/* out[0] = lam'Q lam, out[1] = h'lam after the master; scores[M] for the next sweep */
int rumgpu_cgr_master(rumgpu_cgr* c, int max_outer, double tol, double reg, double* out, float* scores);
/* the masters of `count` runs on the same rows (a batch of bootstrap draws, each draw's round of
 * Bootstrap._solve's per-draw loop, bootstrap.py:308-335) on host threads: run i's results in
 * out[2i..2i+1] and scores[i*M..], equal to rumgpu_cgr_master on that run alone; status[i] its code.
 * ranks (count x n) may be null; else run i with has_rank[i] first adds that ranking's column
 * (rumgpu_cgr_add_ranking). */
int rumgpu_cgr_master_batch(rumgpu_cgr* const* runs, int count, const int32_t* ranks, const int32_t* has_rank,
                            int max_outer, double tol, double reg, double* out, float* scores, int* status);
int rumgpu_cgr_get(rumgpu_cgr* c, int* K, double* lam, double* p_out, uint8_t* columns);

/* ---- Learned warm start: in -> hidden -> hidden -> out MLP (ReLU), bf16 matrix-core forward and
 * backward, Adam on a float32 master copy. No reference counterpart: it maps a bootstrap draw's
 * change of p on the observed cells to a correction of the ADMM state there (the warm start);
 * the float64 certificates still decide every draw. Parameters are one float32 vector: W1 (in x
 * hidden, row-major), b1, W2 (hidden x hidden), b2, W3 (hidden x out), b3. ---- */
int rumgpu_mlp_create(rumgpu_ctx* ctx, int in_dim, int hidden, int out_dim, int max_batch, rumgpu_mlp** out);
int rumgpu_mlp_destroy(rumgpu_mlp* mlp);
int rumgpu_mlp_param_count(rumgpu_mlp* mlp, int64_t* count);
int rumgpu_mlp_set_params(rumgpu_mlp* mlp, const float* theta);
int rumgpu_mlp_get_params(rumgpu_mlp* mlp, float* theta);
/* Y (N x out) = MLP(X (N x in)), row-major host arrays, N <= max_batch */
int rumgpu_mlp_forward(rumgpu_mlp* mlp, int N, const float* X, float* Y);
/* loss = (1/N) sum_i 1/2 ||Y_i - T_i||^2 and its gradient (kept for rumgpu_mlp_adam; copied to
 * grad, param_count floats, when grad is not NULL) */
int rumgpu_mlp_backward(rumgpu_mlp* mlp, int N, const float* X, const float* T, float* loss, float* grad);
/* one Adam step with the kept gradient (bias-corrected moments, step count inside) */
int rumgpu_mlp_adam(rumgpu_mlp* mlp, float lr, float beta1, float beta2, float eps);
/* C (M x N) = A (M x K) B (K x N) through the same bf16 matrix-core tile (a test surface) */
int rumgpu_mlp_gemm_test(rumgpu_ctx* ctx, int M, int N, int K, const float* A, const float* B, float* C);

#ifdef __cplusplus
}
#endif
#endif
