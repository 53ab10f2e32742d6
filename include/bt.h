/*
 * bt.h -- C ABI of the MI355X (gfx950) noisy-circuit engine (libbt.so).
 *
 * Drop-in boundary for BlueTangle.jl's apply path.  Every entry point takes
 * plain pointers and sizes; no C++ types or exceptions cross the ABI.
 *
 * Conventions
 *  - Bit positions are 0-based positions in the basis index.  BlueTangle's
 *    qubit q (1-based, qubit 1 = most significant) is bit (N - q)
 *    (src/bit.jl:9-15).  Callers convert.
 *  - A k-bit matrix is 2^k x 2^k, column-major, complex128 interleaved
 *    (re, im) -- Julia's Matrix{ComplexF64} memory as is.  bits[0] is the
 *    most significant bit of the matrix's local index, i.e. the first tensor
 *    factor acts on bits[0] (BlueTangle's `qubit`; hilbert.jl:30-36).
 *  - Density matrices are stored column-major: entry (r, c) at r + c*2^N,
 *    i.e. vec(rho) is a 2N-bit vector with row bit b at b and column bit b
 *    at b + N.
 *  - Uniform draws: u(seed, traj, i) = SplitMix64-hash counter stream, or
 *    caller-provided arrays; draws are consumed in BlueTangle's order (one per
 *    sampled channel, one per Born measurement, one extra for the MR basis).
 *  - Status: 0 = BT_OK, < 0 = error; bt_last_error() gives a thread-local
 *    message.  A state handle is bound to one HIP stream and is not
 *    thread-safe; distinct handles may be used from distinct host threads.
 */
#ifndef BT_H
#define BT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BT_OK 0
#define BT_ERR_ARG (-1)
#define BT_ERR_UNSUPPORTED (-2)
#define BT_ERR_HIP (-3)
#define BT_ERR_NOMEM (-4)
#define BT_ERR_DRAWS (-5)
#define BT_ERR_STATE (-6)
#define BT_ERR_BUDGET (-7)
#define BT_ERR_NODEVICE (-8)

#define BT_SV 0
#define BT_DM 1
#define BT_C64 0
#define BT_C128 1

#define BT_BASIS_Z 0
#define BT_BASIS_X 1
#define BT_BASIS_Y 2
#define BT_BASIS_R 3

typedef struct bt_state_s* bt_state;
typedef struct bt_circuit_s* bt_circuit;
typedef struct bt_server_s* bt_server;
typedef struct bt_graph_s* bt_graph;

/* ---- library ------------------------------------------------------------ */
const char* bt_last_error(void);
int bt_version(void);
/* Select and initialise a HIP device (BlueTangle has no device notion). */
int bt_init(int device);
/* Counter-based uniform stream shared by the host shim and the kernels. */
double bt_uniform(uint64_t seed, uint64_t traj, uint64_t draw);

/* ---- states (replace zero_state / SparseVector / SparseMatrixCSC,
 *      hilbert.jl:866, ops.jl:809-810) ---------------------------------- */
int bt_state_create(int nqubits, int kind, int prec, bt_state* out);
int bt_state_destroy(bt_state s);
int bt_state_info(bt_state s, int* nqubits, int* kind, int* prec);
/* |index> (SV) or |index><index| (DM); zero_state is index 0 */
int bt_state_set_basis(bt_state s, uint64_t index);
int bt_state_copy(bt_state dst, bt_state src);
/* n = number of complex entries (2^N for SV, 4^N for DM), complex128 interleaved */
int bt_state_upload(bt_state s, const double* data, uint64_t n);
int bt_state_download(bt_state s, double* data, uint64_t n);
/* device pointer and HIP stream of a state, for zero-copy interop */
int bt_state_device_ptr(bt_state s, void** ptr, void** stream);
/* Wait for the work queued on the state's stream. */
int bt_state_sync(bt_state s);
/* Device memory (bytes) a state of this shape would hold. */
uint64_t bt_state_bytes(int nqubits, int kind, int prec);

/* ---- per-op entry points (apply(state, op; noise), hilbert.jl:469-515) -- */
/* Gate: op.expand(N) * state (hilbert.jl:505, hilbert.jl:18-159).
 * SV: psi <- M psi on bits (k <= 5) where all bits of ctrl_mask are 1.
 * DM: rho <- E rho E' (hilbert.jl:655-656), k <= 3. */
int bt_apply_matrix(bt_state s, int k, const double* mat, const int* bits, uint64_t ctrl_mask);
/* OpF(name, mat) (struct.jl:716-720): psi <- M psi with a full 2^N x 2^N
 * matrix, column-major complex128; FP64 accumulation.  SV states only (the
 * reference's matrix OpF has no density-matrix method). */
int bt_apply_full(bt_state s, const double* mat);
/* Reduced density matrix over bits (k <= 3), FP64, column-major 2^k x 2^k;
 * bits[0] = MSB (partial_trace, linalg.jl:83-230).  SV states only. */
int bt_reduced_rho(bt_state s, int k, const int* bits, double* rho_out, double* norm2_out);
/* Trajectory Kraus channel (__QuantumChannel_new_apply, struct.jl:31-55):
 * probabilities p_i = Re tr(K_i rho_A K_i') with rho_A over prob_bits
 * (ascending qubit order, the reference's partial_trace order), draw
 * `u <= cumsum` (hilbert.jl:810-819), apply K at apply_bits, renormalise.
 * chosen / probs_out may be NULL. */
int bt_apply_kraus_traj(bt_state s, int k, int nkraus, const double* kraus, const int* apply_bits,
                        const int* prob_bits, double u, int* chosen, double* probs_out);
/* Density-matrix channel rho <- sum_i K_i rho K_i' (struct.jl:58-76), k <= 2. */
int bt_apply_kraus_dm(bt_state s, int k, int nkraus, const double* kraus, const int* bits);
/* Born measurement with basis rotation (_born_measure / born_measure_Z,
 * hilbert.jl:669-696; MR basis pick, struct.jl:565-571).  u[0] is the
 * outcome draw, or for BT_BASIS_R u[0] picks [X,Y,Z][floor(3u)] and u[1] is
 * the outcome draw.  outcome = u < prob0 ? 0 : 1.  SV states only. */
int bt_measure(bt_state s, int bit, int basis, const double* u, int* outcome, double* prob0);
/* _reset_Z (hilbert.jl:752-759): Z measurement then X on outcome 1. */
int bt_reset(bt_state s, int bit, double u, int* outcome);
/* |psi_i|^2 or diag(rho) in FP64 (sample_exact, ops.jl:98-132) */
int bt_probabilities(bt_state s, double* out);
int bt_norm2(bt_state s, double* out);
/* <a|b> = sum conj(a_i) b_i in FP64, either precision on either side (same
 * shape): fidelity |<a|b>|^2 / (<a|a> <b|b>) of a complex64 result against
 * its complex128 counterpart. */
int bt_inner(bt_state a, bt_state b, double* re, double* im);
/* Expectation value of O = P(xmask, zmask) * G, identity elsewhere, in FP64:
 * <psi|O|psi> on a state vector, tr(rho O) on a density matrix (complex; the
 * reference reports the real part).  P(x, z)[a, b] = (-1)^popcount(a & z) when
 * b = a ^ x: bit b of x flips (X), of z signs (Z); Y on bit b is x and z with a
 * factor -i for the caller.  G is a dense 2^k x 2^k complex128 column-major
 * block on bits[0..k) (bits[0] = MSB, k <= 3, disjoint from x | z; mat may be
 * null when k = 0).  Replaces expect(state|rho, op | op_str | matrix) and
 * correlation(state|rho, op_str, qubits) (func.jl:91-100,142-151), whose
 * operators expand_multi_op / op.expand build (ops.jl:928-944, struct.jl:476-548),
 * and correlation(state, qubits) (func.jl:17-23: a Z string). */
int bt_expect(bt_state s, int k, const double* mat, const int* bits, uint64_t xmask, uint64_t zmask, double* re,
              double* im);
/* Shot sampling of a state vector: shot s returns the smallest basis index i
 * with u[s] * T < C_i, C_i the FP64 cumulative sum of |psi_j|^2 over j <= i in
 * index order and T the total (the state need not be normalised).  u: shots
 * uniforms in [0, 1); out: shots indices.  nboundary (nullable): shots whose
 * u T lies within 1e-12 T of a cumulative boundary, where summation order may
 * decide.  Replaces sample(state, shots) (ops.jl:46-60; StatsBase.sample
 * with Weights, whose RNG the explicit uniforms stand in for). */
int bt_sample(bt_state s, int64_t shots, const double* u, uint64_t* out, int64_t* nboundary);
/* bt_sample without any allocation: the same kernels, results and errors,
 * with every device buffer (chunk sums, prefix, uniforms, indices, flags) in
 * one static 32 MiB device array per GPU (a 2^30-amplitude state with about
 * 1.4 million shots fits; larger requests return BT_ERR_ARG).  No counterpart
 * in the reference: for callers that cannot allocate (tight serving loops,
 * harnesses that forbid hipMalloc).  Calls on one GPU share the array: calls
 * that overlap on one GPU (states on different streams) race. */
int bt_sample_noalloc(bt_state s, int64_t shots, const double* u, uint64_t* out, int64_t* nboundary);
/* DM: FP64 trace and purity sum |rho_ij|^2 */
int bt_trace_purity(bt_state s, double* trace, double* purity);
/* Amplitudes (or rho entries) at n indices, complex128 interleaved out. */
int bt_gather(bt_state s, const uint64_t* idx, uint64_t n, double* out);

/* ---- circuit IR (the op list of run / to_state / to_rho, ops.jl:659-844) */
#define BT_OP_GATE 0
#define BT_OP_CHANNEL 1
#define BT_OP_MEASURE 2
#define BT_OP_RESET 3
#define BT_OP_IF 4

typedef struct {
  int32_t kind;         /* BT_OP_* */
  int32_t k;            /* target bits: gates 1..5, channels 1..3, measure/reset/if 1 */
  int32_t bits[5];      /* target bit positions, bits[0] = MSB of the local index */
  int32_t prob_bits[3]; /* channels: bits of rho_A for the probabilities, MSB first */
  uint64_t ctrl_mask;   /* gates: bits that must be 1 */
  int64_t mat;          /* offset (complex entries) into the matrix pool */
  int32_t nkraus;       /* channels */
  int32_t basis;        /* measure / if: BT_BASIS_* */
  int32_t record;       /* measure / if: 1 = outcome is tracked */
  int32_t br0_begin, br0_end, br1_begin, br1_end; /* if: branch op ranges */
} bt_ir_op;

/* ops[0..ntop) is the top level; branch ops live anywhere in ops[]. */
int bt_circuit_create(int nqubits, const bt_ir_op* ops, int nops, int ntop, const double* pool, int64_t npool,
                      bt_circuit* out);
int bt_circuit_destroy(bt_circuit c);
/* Upper bound of uniform draws one trajectory of the circuit consumes. */
int64_t bt_circuit_max_draws(bt_circuit c);

typedef struct {
  int64_t passes;           /* HBM passes that wrote the state */
  int64_t reduce_passes;    /* read-only reduction passes */
  double bytes;             /* HBM bytes of the passes actually executed */
  int64_t boundary_draws;   /* draws within 1e-6 (c64) / 1e-12 (c128) of a boundary */
  int64_t draws;            /* uniforms consumed */
  double kernel_ms;         /* device time of the pass kernels (HIP events) */
  int64_t launches;         /* pass-kernel launches (bytes / launches per launch) */
} bt_run_stats;

/* One trajectory of a state-vector circuit on state s (run / to_state,
 * ops.jl:659-692,790-795), the state starting as given.  Uniforms come from
 * `uniforms` (nuni of them) or, if NULL, from u(seed, traj, .).  Outcomes of
 * recorded measurements go to outcomes[0..*nout). */
int bt_run_circuit(bt_state s, bt_circuit c, uint64_t seed, uint64_t traj, const double* uniforms, int64_t nuni,
                   int* outcomes, int64_t outcap, int64_t* nout, bt_run_stats* stats);
/* The same circuit on a density matrix (to_rho, ops.jl:806-844); circuits
 * with measurements are rejected as the reference does (hilbert.jl:772). */
int bt_run_circuit_dm(bt_state rho, bt_circuit c, bt_run_stats* stats);
/* ntraj trajectories from |0..0>, uniforms u(seed, traj0 + t, .), the
 * shot loop of run (ops.jl:671-676).  outcomes: ntraj x outcap, nouts: ntraj.
 * final_states (nullable): ntraj x 2^N complex128 interleaved.
 * Lockstep rounds: every round runs the next pass of all resident
 * trajectories in one launch; their interpreters run on host threads. */
int bt_run_batch(bt_circuit c, int prec, uint64_t seed, int64_t traj0, int64_t ntraj, int* outcomes, int64_t outcap,
                 int64_t* nouts, double* final_states, bt_run_stats* stats);
/* bt_run_batch, then one shot per trajectory: the basis index drawn from its
 * final state by the bt_sample rule with the trajectory's next uniform
 * u(seed, traj0 + t, draws used) -- the noisy branch of measure(circuit, shots)
 * (ops.jl:609-621: to_state, final basis, sample(state, 1) per experiment;
 * the final basis rotation is part of the circuit's ops).  samples: ntraj;
 * sample_boundary (nullable): shots within 1e-12 of a cumulative boundary. */
int bt_run_batch_sample(bt_circuit c, int prec, uint64_t seed, int64_t traj0, int64_t ntraj, int* outcomes,
                        int64_t outcap, int64_t* nouts, uint64_t* samples, int64_t* sample_boundary,
                        bt_run_stats* stats);
/* bt_run_batch keeps its worker threads, stream and the slab of resident
 * states between calls (a repeated batch pays no set-up); this frees the
 * cached slab.  Other allocations of the library that run out of HBM free it
 * too.  No counterpart in the reference (which allocates nothing on a device). */
int bt_release_cache(void);
/* The same contract in ONE persistent launch: each workgroup owns a
 * trajectory, runs its interpreter on the device (lane 0) and takes the next
 * trajectory from an atomic counter. */
int bt_run_batch_persistent(bt_circuit c, int prec, uint64_t seed, int64_t traj0, int64_t ntraj, int* outcomes,
                            int64_t outcap, int64_t* nouts, double* final_states, bt_run_stats* stats);

/* Dense 5-qubit block on FP32 MFMA: psi <- M psi on bits (bits[0] = MSB),
 * M 32 x 32 complex128 column-major; complex64 state vectors with N >= 12.
 * The fused form of hilbert(N, mat, ...) * state (hilbert.jl:18-159, 505)
 * for a block of gates. */
int bt_apply_fused5(bt_state s, const double* mat, const int* bits);
/* A noiseless gate-only circuit (the plain apply loop, hilbert.jl:517-553)
 * fused into dense blocks of at most max_k bits (2..5), one pass per block;
 * complex64 with max_k = 5 runs every block on FP32 MFMA.  norm_every > 0:
 * the FP64 norm after every norm_every blocks goes to norms[0..nnorms).
 * *nblocks: blocks applied. */
int bt_run_fused(bt_state s, bt_circuit c, int max_k, int norm_every, double* norms, int64_t nnorms, int64_t* nblocks,
                 bt_run_stats* stats);
/* Fusion-width autotuner for a noiseless gate-only circuit on states like s
 * (its qubit count and precision): times bt_run_fused at max_k = 3, 4 and 5
 * and the register-tile executor (bt_run_circuit) on a scratch state of the
 * same size, and records the fastest on the circuit; bt_run_fused with
 * max_k = 0 then runs it.  *best_k: 3, 4 or 5, or 0 for the register-tile
 * executor; ms (nullable): the four device times in that order (k = 3, 4, 5,
 * tile executor; < 0 where a candidate does not apply).  The results do not
 * depend on the choice beyond rounding (every executor is checked against the
 * oracle).  No counterpart in the reference. */
int bt_autotune_fused(bt_state s, bt_circuit c, int* best_k, double* ms);

/* A fused noiseless circuit captured as a HIP graph on state s: every pass
 * of bt_run_fused recorded once and replayed by bt_graph_launch (one launch,
 * enqueued on s's stream; bt_state_sync waits).  The graph is bound to s
 * and owns its matrices.  complex64 with max_k = 5, or max_k <= 3. */
int bt_graph_create_fused(bt_state s, bt_circuit c, int max_k, bt_graph* out);
/* to_rho (ops.jl:806-844) captured as a HIP graph on density matrix rho: the
 * density-matrix executor's passes (gates as E rho E^dagger, channels as
 * one superoperator each), whose plan depends on nothing they compute,
 * recorded once; bt_graph_launch replays the circuit on rho (reset rho with
 * bt_state_set_basis first).  Circuits with measurements are rejected as in
 * bt_run_circuit_dm. */
int bt_graph_create_dm(bt_state rho, bt_circuit c, bt_graph* out);
int bt_graph_launch(bt_graph g);
int64_t bt_graph_passes(bt_graph g);
int bt_graph_destroy(bt_graph g);

/* Tile width (bits held in LDS per pass) of the circuit executors for a
 * precision; 0 restores the default (13 for BT_C64, 12 for BT_C128).  Set by
 * the autotuner.  Process-wide. */
int bt_set_tile_bits(int prec, int tile_bits);

/* Tuning of the circuit executors (process-wide, per precision): tile bits,
 * low global bits every tile keeps for coalescing (a pass keeps one more when
 * its operators still fit), tiles per workgroup of the
 * rounds executor; 0 restores a default. */
typedef struct {
  int32_t tile_bits;
  int32_t coalesce_bits;
  int32_t tiles_per_block;
  double ms_per_traj;  /* autotune: device time per trajectory of the winner */
} bt_tuning;
int bt_set_tuning(int prec, int tile_bits, int coalesce_bits, int tiles_per_block);
int bt_get_tuning(int prec, int* tile_bits, int* coalesce_bits, int* tiles_per_block);
/* Autotuner: runs `sample` trajectories of the circuit through bt_run_batch
 * for each candidate tile-bits / coalescing setting, then 4 x `sample` for
 * each tiles-per-workgroup candidate of the winner; keeps (and returns) the
 * fastest. */
int bt_autotune_batch(bt_circuit c, int prec, int64_t sample, uint64_t seed, bt_tuning* best);

/* ---- serving: continuous batching with HBM-budget admission ------------- */
typedef struct {
  int64_t id;
  int32_t status;      /* 0 done, < 0 error code */
  int32_t nout;
  double latency_ms;   /* submit -> completion */
  double queue_ms;     /* submit -> admission */
} bt_result;

int bt_server_create(uint64_t hbm_budget_bytes, int max_slots, bt_server* out);
/* Submit one request: a circuit (kept alive by the server until done),
 * BT_SV or BT_DM, precision, seed/traj of its uniform stream.  Returns a
 * request id in *id. */
int bt_server_submit(bt_server srv, bt_circuit c, int kind, int prec, uint64_t seed, uint64_t traj, int64_t* id);
/* Collect finished requests (non-blocking unless wait != 0). */
int bt_server_poll(bt_server srv, bt_result* results, int cap, int wait, int* n);
/* Result payload of a finished request: outcomes and probabilities
 * (2^N doubles; NULL to skip).  Frees the request's device state. */
int bt_server_fetch(bt_server srv, int64_t id, int* outcomes, int64_t outcap, int64_t* nout, double* probs,
                    uint64_t nprobs);
int bt_server_stats(bt_server srv, int* active, int* queued, uint64_t* hbm_in_use);
/* Pre-size the state pool: allocate `count` buffers for requests of this
 * kind, precision and qubit count on the calling thread, so admitting such
 * requests later allocates nothing on the serving thread.  Stops at the HBM
 * budget; *added (nullable) gets the number of buffers added.  No counterpart
 * in the reference (which has no serving loop). */
int bt_server_reserve(bt_server srv, int kind, int prec, int nqubits, int count, int* added);
int bt_server_destroy(bt_server srv);

#ifdef __cplusplus
}
#endif

#endif /* BT_H */
