/*
 * qradient_amd -- C ABI of the MI355X (gfx950) statevector grad_run library.
 *
 * The reference (frederikwilde/qradient) is pure Python and has no FFI; its
 * boundary is the Python object protocol of the circuit classes.  Each entry
 * point below replaces one method of that protocol (reference file:line in
 * the comment); the Python package qradient_amd binds them with ctypes
 * (INTEGRATION.md shows the binding).
 *
 * Conventions
 *  - every function returns QG_OK (0) or a QG_ERR_* code; qg_last_error()
 *    returns a thread-local message for the last failure on this thread;
 *  - pointers are host pointers owned by the caller unless the name says
 *    "dev"; arrays are C-contiguous;
 *  - complex amplitudes cross the ABI as interleaved (re, im) float64 pairs;
 *  - qubit q is index bit (n-1-q): qubit 0 is the most significant bit
 *    (reference qradient/physical_components/state.py:84-87);
 *  - a qg_ctx and the plans/states made from it are single-threaded objects:
 *    calls on one context must not overlap; separate contexts may be used
 *    from separate threads.
 */
#ifndef QRADIENT_AMD_H
#define QRADIENT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QG_OK 0
#define QG_ERR_INVALID 1     /* bad argument: maps to ValueError in Python */
#define QG_ERR_UNSUPPORTED 2 /* valid request this build does not support */
#define QG_ERR_HIP 3         /* HIP runtime failure */
#define QG_ERR_NOMEM 4       /* device memory exhausted */
#define QG_ERR_PENDING 5     /* qg_poll / qg_wait_any: not finished within the timeout */

#define QG_C128 0 /* complex128 state, FP64 arithmetic ("plain" method) */
#define QG_C64 1  /* complex64 state, FP64 arithmetic and FP64 reductions */

#define QG_ENT_CNOT 0 /* CNOT ladder of state.py:209-251 (odd bonds act first) */
#define QG_ENT_CZ 1   /* nearest-neighbour CZ ladder (diagonal variant) */

#define QG_HISTORY (-1) /* ckpt_every: keep the per-pass history, fail if it does not fit */

#define QG_AXIS_X 0
#define QG_AXIS_Y 1
#define QG_AXIS_Z 2

#define QG_TERM_X 0
#define QG_TERM_Y 1
#define QG_TERM_Z 2
#define QG_TERM_ZZ 3

/* One observable term w * P_i (kinds X, Y, Z) or w * Z_i Z_j (kind ZZ, i < j);
 * the flattened form of the reference's observable dict
 * (qradient/physical_components/observable.py:9-18, 44-73). */
typedef struct qg_term {
  int32_t kind;
  int32_t i;
  int32_t j;
  int32_t pad;
  double w;
} qg_term;

typedef struct qg_ctx qg_ctx;
typedef struct qg_plan qg_plan;
typedef struct qg_state qg_state;

const char* qg_last_error(void);
int qg_version(void);

/* Context: one HIP device and one stream (NULL: the library creates a
 * default-priority stream).  The stream argument is a hipStream_t. */
int qg_ctx_create(int device, void* stream, qg_ctx** out);
int qg_ctx_destroy(qg_ctx* ctx);
int qg_ctx_synchronize(qg_ctx* ctx);

/* Measurement utility, no counterpart in the reference: `reps` copies of
 * `bytes` (a multiple of 16) from the device buffer src to dst on the
 * context's stream by a streaming-copy kernel (16 bytes per lane, four
 * accesses in flight per lane, grid = grid_per_cu x the CU count; nontemporal
 * = 1 uses streaming loads and stores).  *ms_out = device milliseconds per
 * copy (HIP events).  The benchmark's copy ceiling for the pass kernels. */
int qg_stream_copy(qg_ctx* ctx, void* dst, const void* src, size_t bytes, int nontemporal, int grid_per_cu,
                   int reps, double* ms_out);

/* ---- McClean circuit (qradient/circuit_logic/mc_clean.py) ----------------
 * Plan for n qubits and L layers.  ckpt_every = 0 lets the planner keep the
 * per-pass state history when it fits in 75% of the device's total memory
 * and otherwise checkpoint (the smallest period whose states fit that
 * budget); k > 0 forces checkpoints every k layers with psi uncomputed in
 * between; QG_HISTORY forces the history (QG_ERR_NOMEM if it does not fit).
 * The choice depends on the shape and the device only, never on the memory
 * free at the time, so one shape always gives the same result bits;
 * qg_plan_ckpt_every reports it.  Replaces the construction of
 * McClean.__init__ (mc_clean.py:9-25). */
int qg_mcclean_plan(qg_ctx* ctx, int n, int L, int dtype, int entangler, int ckpt_every, qg_plan** out);

/* Device workspace (bytes) a McClean plan of this shape uses for a grad run
 * without ini_state (0 for n <= 13: one workgroup runs the whole circuit),
 * and the memory mode it chooses (*ckpt_chosen = 0: the per-pass history,
 * k > 0: checkpoints every k layers; may be NULL).  The same layout rule as
 * qg_mcclean_plan.  ctx may be NULL when ckpt_every is QG_HISTORY or k > 0
 * (the rule then needs no device).  Used to admit requests by footprint. */
int qg_mcclean_workspace_bytes(qg_ctx* ctx, int n, int L, int dtype, int ckpt_every, size_t* bytes,
                               int* ckpt_chosen);

/* McClean.grad_run (mc_clean.py:47-78).  axes: int32 [L*n] in {0,1,2};
 * angles: [L*n]; terms: the observable; ini_state: NULL (|0...0>) or a host
 * complex128 vector [2*2^n] (the RY(pi/4) layer is still applied, as in the
 * reference).  Writes E and grad[L*n] (layer-major). */
int qg_mcclean_grad_run(qg_plan* plan, const int32_t* axes, const double* angles, const qg_term* terms,
                        int nterms, const double* ini_state, double* E_out, double* grad_out);

/* Allocation-free McClean grad run.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It is the
 * whole-circuit path of qg_mcclean_grad_run (n <= 13, no ini_state, one
 * workgroup, FP64 throughout, so the result does not depend on a dtype) and
 * returns bitwise the same E and grad[L*n] as a plan of that shape, without a
 * plan: the tables and results live in one static device array per GPU of
 * QG_NOALLOC_SCRATCH_BYTES, whose address the library takes at
 * qg_ctx_create.  It calls no hipMalloc, hipFree or other allocating routine.
 * Circuits whose tables exceed the array (about 28 bytes per parameter plus
 * the observable) are rejected with QG_ERR_INVALID, as is n > 13.  Every call
 * on a GPU shares that GPU's array, which keeps the last call's contents:
 * calls that overlap on one GPU (from any context or thread) race. */
#define QG_NOALLOC_SCRATCH_BYTES (256u * 1024u)
int qg_mcclean_grad_run_noalloc(qg_ctx* ctx, int n, int L, int entangler, const int32_t* axes,
                                const double* angles, const qg_term* terms, int nterms, double* E_out,
                                double* grad_out);

/* McClean.run_expec_val with exact_expec_val=True (mc_clean.py:27-45). */
int qg_mcclean_expec_val(qg_plan* plan, const int32_t* axes, const double* angles, const qg_term* terms,
                         int nterms, const double* ini_state, double* E_out);

/* Projector probabilities of parameter-shifted McClean circuits, for the
 * shot-based parameter-shift gradients McClean.sample_grad and
 * McClean.sample_grad_with_component_sampling (mc_clean.py:117-198).  Shift s
 * adds shift_delta[s] to angles[shift_param[s]] (flat index i*n + q; the
 * reference's manual rotation of pi/2 or -pi/2 after layer i, which is the
 * same circuit); prob[s*nterms + k] = |P_k^+ psi_s|^2 of term k in the final
 * state psi_s of that circuit (weights ignored), as qg_plan_term_prob gives
 * it.  The unshifted circuit runs first and keeps its layer-end states
 * (history) or checkpoints; each shifted circuit restarts from the nearest
 * one before its shifted layer, like the reference's restart from
 * state_history[i].  Afterwards the plan's state is the unshifted circuit's
 * final state.  The shots are drawn by the caller.  Needs n <= 31. */
int qg_mcclean_shift_term_prob(qg_plan* plan, const int32_t* axes, const double* angles, const double* ini_state,
                               int nshift, const int32_t* shift_param, const double* shift_delta,
                               const qg_term* terms, int nterms, double* prob);

/* The same shifted circuits, measured in a basis ordered by `order` (int32
 * [2^n], order[k] = the amplitude at position k; NULL: natural order): for
 * shift s and shot m, idx[s*shots + m] = the first position k whose
 * cumulative probability sum_{k' <= k} |psi_s[order[k']]|^2 reaches
 * u[s*shots + m], or 0 when none does.  That is scipy's
 * rv_discrete(values=(arange(2^n), dist)).rvs draw for the uniform u, as
 * McClean.sample_grad_dense samples the observable's eigenbasis
 * (mc_clean.py:207-268): for a diagonal observable, order sorts its diagonal
 * ascending (the order of np.linalg.eigh's eigenvalues) and the drawn
 * eigenvalue is diag[order[idx]].  The uniforms are drawn by the caller. */
int qg_mcclean_shift_sample(qg_plan* plan, const int32_t* axes, const double* angles, const double* ini_state,
                            int nshift, const int32_t* shift_param, const double* shift_delta, const int32_t* order,
                            int shots, const double* u, int32_t* idx);

/* ---- batched McClean: many parameter sets of one shape --------------------
 * The reference runs one circuit per McClean.grad_run call (mc_clean.py:47-78);
 * its callers that need many (an optimiser over several starting points,
 * optimization.py:69-91) loop.  A batch plan holds up to `batch` circuits of n
 * qubits and L layers, each with its own axes and angles, sharing one z / zz
 * observable; every pass is one launch over all circuits resident at once
 * (chunks sized to device memory).  Tile shape, memory mode (the per-pass
 * history) and per-circuit grids are those of a single plan of the shape,
 * so each circuit's result is bitwise what qg_mcclean_grad_run gives it. */
int qg_mcclean_batch_plan(qg_ctx* ctx, int n, int L, int dtype, int entangler, int batch, qg_plan** out);
/* axes int32 [B*L*n], angles [B*L*n] (circuit b at b*L*n); E_out[B],
 * grad_out[B*L*n].  Observables with x / y terms: QG_ERR_UNSUPPORTED. */
int qg_mcclean_batch_grad_run(qg_plan* plan, int B, const int32_t* axes, const double* angles, const qg_term* terms,
                              int nterms, double* E_out, double* grad_out);

/* ---- QAOA circuit (qradient/circuit_logic/qaoa.py) -----------------------
 * Plan for `batch` circuits of n qubits and p layers each. */
int qg_qaoa_plan(qg_ctx* ctx, int n, int p, int dtype, int batch, qg_plan** out);

/* Qaoa.grad_run (qaoa.py:40-70) for B <= batch circuits in one launch
 * sequence.  Circuit b has terms[term_off[b] .. term_off[b+1]) (z / zz kinds:
 * the classical Hamiltonian C of state.py:261-297), betas[b*p..], gammas[b*p..].
 * grad_out[b*p*2 + i*2 + 0] = dE/dbeta_i, [.. + 1] = dE/dgamma_i. */
int qg_qaoa_grad_run(qg_plan* plan, int B, const qg_term* terms, const int32_t* term_off, const double* betas,
                     const double* gammas, double* E_out, double* grad_out);

/* Batch shape of a QAOA plan (no counterpart in the reference): kmax local
 * bits per pass tile (0: the default rule), at least cmin contiguous low bits
 * in high-bit passes, `chunk` circuits resident at once (0: as many as fit in
 * device memory; fewer circuits per chunk keep more of a chunk's states in
 * the 256 MB memory-side cache), and at most wg_per_cu workgroups per CU in
 * a batched pass launch, over all circuits.  The plan is laid out again; used
 * by the autotuner. */
int qg_qaoa_tune(qg_plan* plan, int kmax, int cmin, int chunk, int wg_per_cu);

/* Qaoa.run_expec_val with exact_expec_val=True (qaoa.py:23-38). */
int qg_qaoa_expec_val(qg_plan* plan, int B, const qg_term* terms, const int32_t* term_off, const double* betas,
                      const double* gammas, double* E_out);
/* As qg_qaoa_grad_run / qg_qaoa_expec_val, from given initial states instead
 * of |+>^n (qaoa.py:40-49 with ini_state): ini_states holds B states of 2^n
 * complex128 amplitudes (interleaved re, im), circuit b at b * 2^(n+1)
 * doubles.  The states are read, never modified (the reference multiplies its
 * ini_state in place). */
int qg_qaoa_grad_run_ini(qg_plan* plan, int B, const qg_term* terms, const int32_t* term_off, const double* betas,
                         const double* gammas, const double* ini_states, double* E_out, double* grad_out);
int qg_qaoa_expec_val_ini(qg_plan* plan, int B, const qg_term* terms, const int32_t* term_off, const double* betas,
                          const double* gammas, const double* ini_states, double* E_out);

/* Lays a McClean plan out again for another circuit shape: n qubits, L
 * layers, ckpt_every as for qg_mcclean_plan; dtype, entangler, context and
 * tile options are kept.  The plan's small device tables only grow, so after
 * it has held a shape at least as large in every respect (qubits, layers,
 * passes) the call makes no device allocation.  A workspace supplied with
 * qg_plan_set_workspace stays attached if it is large enough for the new
 * layout and is detached otherwise (supply another, or the next run
 * allocates one); an owned one that is too small is freed.  For serving
 * loops whose requests change shape.  No counterpart in the reference (one
 * circuit object per shape there). */
int qg_plan_reshape(qg_plan* plan, int n, int L, int ckpt_every);

int qg_plan_workspace_size(const qg_plan* plan, size_t* bytes);
/* Hand the plan caller-owned device memory (e.g. a torch tensor) of at least
 * qg_plan_workspace_size bytes; without it the plan allocates on first use. */
int qg_plan_set_workspace(qg_plan* plan, void* dev_ptr, size_t bytes);
/* Copy the final state of the last forward run (complex128 interleaved). */
int qg_plan_get_state(qg_plan* plan, double* out);

/* prob[k] = |P_k^+ psi|^2 = <psi|(I + P_k)/2|psi>, the probability of the +1
 * outcome of term k of terms[0, nterms) (weights ignored), in the state of the
 * plan's last forward run, in the order given.  Replaces the projector
 * probabilities of ParametrizedCircuit.sample_expec_val
 * (qradient/circuit_logic/base.py:22-46; Projector at
 * qradient/physical_components/observable.py:82-104, 126-177); the shots are
 * drawn by the caller.  Needs n <= 31.  After a batched QAOA run (B > 1) the
 * state is that of the first circuit of the last chunk on the pass kernels,
 * and absent after the whole-circuit kernel. */
int qg_plan_term_prob(qg_plan* plan, const qg_term* terms, int nterms, double* prob);
/* E = sum_k w_k <psi|P_k|psi> over terms[0, nterms) in the state of the plan's
 * last forward run: ParametrizedCircuit.expec_val (base.py:17-20) without
 * forming the matrix, for x / y / z / zz terms at any n <= 31.  Used for the
 * full-observable expectation of McClean.grad_run_with_component_sampling
 * (mc_clean.py:80-115).  A grad run with checkpoints overwrites the final
 * state (QG_ERR_INVALID then). */
int qg_plan_expec_val(qg_plan* plan, const qg_term* terms, int nterms, double* E_out);
/* Device time of the last run in milliseconds (hipEvents around the launches). */
int qg_plan_last_ms(const qg_plan* plan, double* ms);
/* Record a HIP event before every kernel launch of later runs (on = 1).  Each
 * call starts a new series of profiled runs; a plan keeps the latest 256. */
int qg_plan_set_profiling(qg_plan* plan, int on);
/* Per-launch device times (ms) of the profiled runs of the current series,
 * run after run, and their kind (0 forward pass, 1 reverse two-state pass,
 * 2 forward pass fused with the observable, 3 another launch); *count
 * receives the launch count, at most cap are written.  The times are read
 * from the events here, so profiled runs pay only for the event records. */
int qg_plan_pass_times(const qg_plan* plan, int cap, float* ms, int* kinds, int* count);
/* Tile shape of a McClean plan: kmax local bits per pass (4..14; 14 only for
 * QG_C64 plans that keep the per-pass history with FP32 tiles), at least
 * cmin contiguous low bits in high-bit passes, and log2 amplitudes per thread
 * of the forward and adjoint passes (r_fwd = r_rev = 4).  The plan is laid out
 * again (its pass count, memory mode and workspace size may change; a
 * caller-supplied workspace that is too small is an error).  Used by the
 * autotuner.  Default: the fewest passes per layer that 13 local bits allow,
 * with the smallest kmax (>= 11 for QG_C64, >= 12 for QG_C128) that keeps it;
 * kmax 13 for QG_C64 states of 4 GiB and more (n >= 29). */
int qg_plan_tune(qg_plan* plan, int kmax, int cmin, int r_fwd, int r_rev);
/* on = 1 (default): circuits with n <= kmax run whole in one workgroup each
 * (one launch per call, FP64 throughout); on = 0 uses the pass kernels. */
int qg_plan_set_fused(qg_plan* plan, int on);
/* on = 1: a run's pass launches are captured once as a HIP graph and
 * replayed while the plan's launch list is unchanged (a run uploads its angles
 * and tables to the plan's descriptor table before the replay); on = 0
 * (default) launches every pass directly.  Profiled runs launch directly.
 * Enable it only where no other thread makes device-synchronising HIP calls
 * (allocation, synchronous copies) while this plan runs: those invalidate a
 * capture in progress. */
int qg_plan_set_graph(qg_plan* plan, int on);
/* LDS tile precision of a QG_C64 McClean plan: mode 1 keeps the pass tiles
 * in FP64 (the state is rounded to complex64 once per pass, when stored),
 * 0 in FP32 (also rounded at each of a pass's LDS transposes; faster), 2
 * FP64 in the forward passes and FP32 in the adjoint passes, -1 (default)
 * picks 2 when the circuit has more than 200 passes, else 0.  The plan is
 * laid out again.  qg_plan_lds64 returns the mode in effect (0, 1 or 2).
 * No counterpart in the reference (complex128 only). */
int qg_plan_set_lds64(qg_plan* plan, int mode);
int qg_plan_lds64(const qg_plan* plan, int* out);
/* The checkpoint period a McClean plan chose (0: full per-pass history). */
int qg_plan_ckpt_every(qg_plan* plan, int* out);
int qg_plan_destroy(qg_plan* plan);

/* ---- device-resident State (qradient/physical_components/state.py) --------
 * ini: '0' or '+' (state.py:61-71). */
int qg_state_create(qg_ctx* ctx, int n, int dtype, char ini, qg_state** out);
int qg_state_destroy(qg_state* st);
int qg_state_reset(qg_state* st, char ini);                        /* state.py:61-71 */
int qg_state_set(qg_state* st, const double* vec);                 /* State.vec = ...   */
int qg_state_get(qg_state* st, double* vec);                       /* ... = State.vec   */
/* rotation (deriv = 0) or its angle derivative (deriv = 1) about axis on qubit q:
 * xrot/yrot/zrot state.py:90-92,142-144,168-170; dxrot/dyrot/dzrot 94-97,146-149,172-175 */
int qg_state_rot(qg_state* st, int axis, double angle, int q, int deriv);
int qg_state_rot_all(qg_state* st, int axis, double angle);        /* Qaoa.__xrot_all, qaoa.py:192-194 */
int qg_state_cnot_ladder(qg_state* st, int stacking);              /* state.py:243-251 */
int qg_state_cz_ladder(qg_state* st);
int qg_state_cnot(qg_state* st, int control, int target);          /* state.py:198-199, 336-356 */
/* vec <- exp(-i*angle*D) vec with D the z/zz part of terms (state.py:299-301) */
int qg_state_exp_diag(qg_state* st, const qg_term* terms, int nterms, double angle);
/* vec <- O vec for the full observable (multiply_matrix, mc_clean.py:65) */
int qg_state_multiply_observable(qg_state* st, const qg_term* terms, int nterms);
/* <vec|O|vec> in FP64 (ParametrizedCircuit.expec_val, base.py:17-20) */
int qg_state_expec_val(qg_state* st, const qg_term* terms, int nterms, double* out);
/* vec <- (re + i im) vec (the -1j factors of state.py:121-123, 319-321) */
int qg_state_scale(qg_state* st, double re, double im);
/* 1 - ||vec|| (state.py:331-332) */
int qg_state_norm_error(qg_state* st, double* out);

/* prob[k] = |P_k^+ vec|^2 for each term (the projector probabilities of
 * base.py:22-46), as qg_plan_term_prob. */
int qg_state_term_prob(qg_state* st, const qg_term* terms, int nterms, double* prob);
/* Basis draws from the state, as qg_mcclean_shift_sample draws them from a
 * shifted circuit's final state: the inverse CDF of |v|^2 over the positions
 * of `order` (NULL: natural order) at the given uniforms (Qaoa.__sample,
 * qaoa.py:196-198, with the classical Hamiltonian's values in natural order
 * as eigenvalues).  Needs n <= 31. */
int qg_state_sample_basis(qg_state* st, const int32_t* order, int shots, const double* u, int32_t* idx);
/* dst.vec = src.vec on the device (same n and dtype): the reference's
 * state.vec[:] = state_history[k] (qaoa.py:139, 151) without a host copy. */
int qg_state_copy(qg_state* dst, const qg_state* src);

/* ---- serving: continuous batching of McClean grad runs ---------------------
 * No counterpart in the reference, whose caller of grad_run is the optimiser
 * loop (optimization.py:69-91), one circuit at a time.  A server owns one
 * device arena of arena_bytes (allocated at creation) and `slots` worker
 * threads, each with its own stream and one plan re-laid out per request
 * (qg_plan_reshape), so requests of different shapes overlap on the device
 * and a shape change allocates nothing once the plans have grown.  A request
 * is admitted when a worker is idle and the arena can hold its workspace
 * (qg_server_footprint); requests that fit may overtake a head-of-queue
 * request that does not at most backfill_limit times (with
 * QG_SERVER_SHORTEST_FIRST: a request may be overtaken by later arrivals at
 * most backfill_limit times).  A request's memory
 * mode is a function of its shape only: the per-pass history when that
 * needs at most history_max_bytes, else checkpoints every ckpt_every layers,
 * so its result is bitwise the same alone and in any mix.  The observable is
 * any set of terms; the entangler is the reference's CNOT ladder.  All
 * server functions may be called from any thread. */
#define QG_SERVER_SIMULATE 1 /* flags: no device; a request sleeps ~1 us per MiB of
                                workspace and returns E = n, grad = L (scheduler tests) */
#define QG_SERVER_SHORTEST_FIRST 2 /* flags: admit the cheapest queued request that fits
                                      (estimated work (6 L + 3) 2^n) instead of the oldest;
                                      a request overtaken backfill_limit times by later
                                      arrivals is served next, in arrival order */
typedef struct qg_server qg_server;
typedef struct qg_request_info {
  double queue_ms;        /* submit -> start on a worker (host clock) */
  double run_ms;          /* start -> results ready */
  double latency_ms;      /* submit -> results ready */
  double device_start_ms; /* device time of the request's first launch, ms after server creation */
  double device_end_ms;   /* device time after its last launch */
  int32_t slot;           /* worker that ran it */
  int32_t ckpt_every;     /* memory mode: 0 per-pass history, k > 0 checkpoints every k layers */
  uint64_t footprint;     /* workspace bytes it held in the arena */
} qg_request_info;
typedef struct qg_server_counters {
  int64_t queued, running, completed;
  uint64_t arena_bytes, arena_in_use, arena_peak;
} qg_server_counters;

int qg_server_create(int device, int slots, int dtype, size_t arena_bytes, size_t history_max_bytes, int ckpt_every,
                     int backfill_limit, int flags, qg_server** out);
/* Workspace bytes a request of this shape holds in the arena and its memory
 * mode (*ckpt_every = 0: history, k > 0: checkpoints; may be NULL). */
int qg_server_footprint(qg_server* server, int n, int L, size_t* bytes, int* ckpt_every);
/* Queue McClean.grad_run (mc_clean.py:47-78) of one circuit: axes int32 and
 * angles [L*n], the observable terms.  The inputs are copied.  Invalid
 * inputs fail here (QG_ERR_INVALID); a workspace larger than the arena fails
 * with QG_ERR_NOMEM.  *ticket identifies the request. */
int qg_submit(qg_server* server, int n, int L, const int32_t* axes, const double* angles, const qg_term* terms,
              int nterms, int64_t* ticket);
/* Waits up to timeout_ms (< 0: no limit) for the request; when it has
 * finished, writes E, grad[L*n] and info (each may be NULL), forgets the
 * ticket and returns the request's status; QG_ERR_PENDING if it is still
 * queued or running. */
int qg_poll(qg_server* server, int64_t ticket, int timeout_ms, double* E_out, double* grad_out,
            qg_request_info* info);
/* Waits up to timeout_ms (< 0: no limit) for any request to finish that no
 * earlier qg_wait_any returned; *ticket receives it (collect it with
 * qg_poll).  QG_ERR_PENDING on timeout. */
int qg_wait_any(qg_server* server, int timeout_ms, int64_t* ticket);
int qg_server_stats(qg_server* server, qg_server_counters* out);
/* Runs every queued request to completion, then frees the server. */
int qg_server_destroy(qg_server* server);

#ifdef __cplusplus
}
#endif

#endif /* QRADIENT_AMD_H */
