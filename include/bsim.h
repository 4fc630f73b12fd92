/*
 * bsim.h -- C ABI of libbsim, the MI355X (gfx950) engine behind
 * BraketSimulator.jl's evolution hot path.
 *
 * Conventions shared by every entry point
 * ---------------------------------------
 *  - Amplitudes are complex128, interleaved (re, im): byte-identical to
 *    Julia's Vector{ComplexF64}.  A density matrix is column-major N x N.
 *  - A state handle owns its device buffer and one HIP stream.  Calls on one
 *    handle are ordered on that stream; a handle is not thread-safe, calls on
 *    different handles may run concurrently (as the reference's batch pool
 *    gives each worker its own simulator, src/BraketSimulator.jl:387).
 *  - "Vector qubits": a gate entry point acts on the state viewed as a vector
 *    of 2^vq amplitudes -- vq = n for a state vector and vq = 2n for a
 *    density matrix, exactly as the reference reshapes rho to a 4^n vector
 *    before calling apply_gate! (src/dm_simulator.jl:135-138).  Qubit q of
 *    that vector is amplitude-index bit vq-1-q (src/gate_kernels.jl:193).
 *  - Matrices are column-major interleaved complex128 (the memory of a Julia
 *    SMatrix{..,ComplexF64}), `conj` != 0 applies conj(M) (the Val{true}
 *    dispatch, src/gate_kernels.jl:675-680).
 *  - Every function returns BSIM_OK (0) or a negative code; the message of
 *    the last failure on the calling thread is read with bsim_last_error.
 *  - Host buffers are caller-owned and read/written synchronously.
 */
#ifndef BSIM_H
#define BSIM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BSIM_OK 0
#define BSIM_EINVAL -1   /* bad argument (target out of range, shape) */
#define BSIM_ENOMEM -2   /* device allocation failed */
#define BSIM_EHIP -3     /* HIP runtime error */
#define BSIM_ENOTSUP -4  /* outside the supported range */
#define BSIM_ENODEV -5   /* no usable gfx950 device */

#define BSIM_SV 0
#define BSIM_DM 1

typedef struct bsim_state bsim_state;

/* ---- library ------------------------------------------------------ */

/* Version of the ABI (major*10000 + minor*100 + patch): 10100. */
int bsim_version(void);
/* Copy the last error message of this thread into buf (NUL-terminated);
 * returns its full length. */
int bsim_last_error(char* buf, int64_t len);
/* Device name, CU count and HBM size of the current device. */
int bsim_device_info(char* name, int64_t name_len, int* cu_count, int64_t* hbm_bytes);

/* ---- state (src/sv_simulator.jl:6-127, src/dm_simulator.jl:6-128) ---- */

/* StateVectorSimulator / DensityMatrixSimulator storage + init():
 * |0...0> or |0...0><0...0| (src/sv_simulator.jl:43-51, dm_simulator.jl:41-49). */
int bsim_state_create(int qubit_count, int kind, bsim_state** out);
int bsim_state_destroy(bsim_state* h);
/* reinit! (src/sv_simulator.jl:96-127, src/dm_simulator.jl:96-128):
 * back to the all-zero state, keeping the allocation. */
int bsim_state_reset(bsim_state* h);
int bsim_state_qubits(const bsim_state* h, int* qubit_count, int* kind);
/* Copy n_complex amplitudes from / to host memory (state_vector(),
 * density_matrix(): src/sv_simulator.jl:150, src/dm_simulator.jl:261). */
int bsim_state_upload(bsim_state* h, const double* host, int64_t n_complex);
int bsim_state_download(bsim_state* h, double* host, int64_t n_complex);
/* copyto!(dst, src) (src/sv_simulator.jl:88-94). */
int bsim_state_copy(bsim_state* dst, const bsim_state* src);
/* Block until every call issued on h has finished. */
int bsim_synchronize(bsim_state* h);
/* Device address of the amplitudes (for zero-copy interop). */
int bsim_state_device_ptr(bsim_state* h, void** dptr);

/* ---- per-gate kernels (src/gate_kernels.jl) ---------------------- */

/* apply_gate!(::SMatrix{2,2}, sv, q): src/gate_kernels.jl:489-537. */
int bsim_apply_1q(bsim_state* h, int q, const double* m, int conj);
/* apply_gate!(::SMatrix{4,4}, sv::Vector, t1, t2): src/gate_kernels.jl:540-576.
 * The reference kernel gathers (00, t1-flipped, t2-flipped, 11): t1 is bit 0
 * of the 4x4 index. */
int bsim_apply_2q(bsim_state* h, int t1, int t2, const double* m, int conj);
/* apply_controlled_gate! for every arity: 1 control/1 target
 * (src/gate_kernels.jl:579-606), 1 control/2 targets (:608-636, t1 = bit 0),
 * 2 controls/1 target (:638-670).  cvals[i] is the bit value control i
 * selects (Control{G,B}.bitvals, CNot & co use 1).  Generalised to nc <= 8
 * controls and nt <= 2 targets. */
int bsim_apply_controlled(bsim_state* h, int nc, const int* controls, const int* cvals, int nt,
                          const int* targets, const double* m, int conj);
/* k-qubit dense fallback apply_gate!(::SMatrix{N,N}, sv, targets...):
 * src/gate_kernels.jl:753-775.  Bit j of the matrix index is the j-th
 * smallest endian target: the given order is ignored.  k <= 12 (4..6 on the
 * FP64 matrix cores, 7..12 through LDS-staged groups). */
int bsim_apply_kq(bsim_state* h, int k, const int* targets, const double* m, int conj);
/* Diagonal gates (Diagonal{SVector} through the kernels above and
 * custom_gates.jl:364-388 for MultiRZ / GPhase): amplitude i is multiplied
 * by d[j], where bit b of j is the amplitude bit of qubit targets[b].
 * k <= 8. */
int bsim_apply_diagonal(bsim_state* h, int k, const int* targets, const double* d, int conj);
/* DoubleExcitation kernel: src/custom_gates.jl:390-416 (angle * pow). */
int bsim_apply_double_excitation(bsim_state* h, const int* targets4, double angle);

/* ---- instruction stream, fused evolution (src/sv_simulator.jl:140-148,
 *      src/dm_simulator.jl:170-179) ------------------------------------ */

#define BSIM_OP_GATE 0      /* dense 2^k x 2^k unitary; targets[b] = qubit of matrix bit b */
#define BSIM_OP_DIAG 1      /* diagonal gate, 2^k entries, same target meaning */
#define BSIM_OP_KRAUS_REF 2 /* DM only: Kraus channel, reference block-diagonal semantics */
#define BSIM_OP_KRAUS_PHYS 3 /* DM only: Kraus channel on every block */
#define BSIM_OP_PHASEFLIP 4 /* DM only: PhaseFlip(p) elementwise, matrix[0] = p */
#define BSIM_OP_SUPEROP 5   /* DM only: 4^k x 4^k superoperator, k <= 5 (bsim_dm_apply_superop) */

#define BSIM_MAX_TARGETS 8
#define BSIM_MAX_CONTROLS 8

/* One instruction.  Qubit numbers are reference qubits of the simulator
 * (0..n-1); for a density matrix the library applies the gate as
 * U on the column bits and conj(U) on the row bits. */
typedef struct bsim_op {
    int32_t type;
    int32_t n_targets;
    int32_t n_controls;
    int32_t n_kraus;                       /* Kraus ops: number of matrices */
    int32_t targets[BSIM_MAX_TARGETS];     /* targets[b] <-> matrix index bit b */
    int32_t controls[BSIM_MAX_CONTROLS];
    int32_t control_values[BSIM_MAX_CONTROLS];
    const double* matrix;                  /* gate: 4^k complex; diag: 2^k; Kraus: n_kraus * 4^k */
    /* differentiation: up to 3 parameters per gate, each with dU/dtheta */
    int32_t n_params;
    int32_t param_index[3];
    const double* dmatrix[3];
} bsim_op;

/* evolve!: apply ops in order.  mode 0 = fused tiles (default), 1 = one
 * launch per op (the reference's loop shape).  Results are the same. */
int bsim_evolve(bsim_state* h, const bsim_op* ops, int n_ops, int mode);

/* ---- compact op-list image ------------------------------------------
 * The reference builds its instruction stream from OpenQASM 3 in Julia
 * (to_circuit, src/circuit.jl:277-310) and hands it to evolve! as objects.
 * A bsim_op array holds host pointers; its packed image is self-contained
 * bytes (no pointers) that can be stored, sent to another process and
 * evolved, captured or batched from there.  Layout, little-endian, 8-byte
 * aligned:
 *   bsim_packed_header;
 *   n_ops x bsim_packed_op  (matrix / dmatrix: index of the first complex
 *                            value in the value array, -1 when absent);
 *   n_values x complex128   (re, im) -- per op its matrix (gate 4^k, diag
 *                            2^k, Kraus n_kraus 4^k, PhaseFlip 1 (p, 0),
 *                            superoperator 16^k values), then its
 *                            n_params derivative matrices (4^k each). */
#define BSIM_PACK_VERSION 1
typedef struct bsim_packed_header {
    char magic[4];          /* "BSOP" */
    uint32_t version;       /* BSIM_PACK_VERSION */
    uint32_t n_ops;
    uint32_t record_bytes;  /* sizeof(bsim_packed_op) */
    uint64_t n_values;
} bsim_packed_header;
typedef struct bsim_packed_op {
    int32_t type, n_targets, n_controls, n_kraus;
    int32_t targets[BSIM_MAX_TARGETS];
    int32_t controls[BSIM_MAX_CONTROLS];
    int32_t control_values[BSIM_MAX_CONTROLS];
    int64_t matrix;
    int32_t n_params;
    int32_t param_index[3];
    int64_t dmatrix[3];
} bsim_packed_op;
/* Image of ops[0..n_ops) into buf (cap bytes; 8-byte aligned); *size gets
 * the bytes needed (buf = NULL: size only).  BSIM_EINVAL when cap is
 * smaller or an op is malformed (null matrix, counts out of range). */
int bsim_ops_pack(const bsim_op* ops, int n_ops, void* buf, size_t cap, size_t* size);
/* Validates an image of `size` bytes (magic, version, record size, every op's
 * counts and value ranges inside the buffer) and fills ops[0..*n_ops) with
 * pointers into buf -- no copy: the ops are valid while buf lives unchanged.
 * ops = NULL: *n_ops only; BSIM_EINVAL when cap < *n_ops or the image is
 * malformed.  Qubit ranges are checked by the call that takes the ops. */
int bsim_ops_unpack(const void* buf, size_t size, bsim_op* ops, int cap, int* n_ops);
/* bsim_evolve of an image. */
int bsim_evolve_packed(bsim_state* h, const void* buf, size_t size, int mode);

/* evolve! (fused mode) without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc).  The plan's tables live in a
 * fixed 4 MiB device array per GPU instead of the handle's workspace; plans
 * whose tables exceed it, or that hold ops outside the tile passes (dense
 * gates on 3+ targets, Kraus channels run on their own), return
 * BSIM_ENOTSUP.  The array is shared by every call on a GPU: calls that
 * overlap in time on one GPU race.  Returns after the evolution finished. */
int bsim_evolve_noalloc(bsim_state* h, const bsim_op* ops, int n_ops);

/* Planner statistics for the fused path, computed on the host only:
 * stats[0] = HBM passes, stats[1] = per-op launches outside tiles,
 * stats[2] = LDS sweeps, stats[3] = ops fused. */
int bsim_plan_stats(int vector_qubits, int is_dm, const bsim_op* ops, int n_ops, int64_t* stats);

/* Tile configuration of the fused path (tile_bits 6..13, contiguous low
 * bits 0..tile_bits-2); 0 keeps the current value. */
int bsim_set_tile_config(int tile_bits, int low_bits);
int bsim_get_tile_config(int* tile_bits, int* low_bits);
/* Register bits per LDS sweep of the fused path: 4 (16-amplitude
 * sub-cubes, 256-thread workgroups; the default) or 3 (8 amplitudes, 512
 * threads, four waves per SIMD); 0 keeps the current value, *old (may be
 * NULL) receives the previous one.  Passes holding 4-bit ops or reference
 * Kraus ops (density matrices) always use 4. */
int bsim_set_sweep_bits(int bits, int* old);
/* Fusion width of the fused path: FP64 operations per amplitude one pass
 * may spend (1-qubit gate 8, 2-qubit 32, diagonal 4, CZ-type sign 1);
 * 0 = unlimited.  Returns the previous value in *old (may be NULL). */
int bsim_set_fusion_limit(int max_cost, int* old);
/* Width of the merged dense blocks of the fused path: on vectors of 20+
 * bits, runs of adjacent dense ops around a 3+-target op are multiplied
 * into one op of at most `bits` bits (6 by default, 5 below 22 bits) that
 * runs on the FP64 matrix cores; 0 = no merging, -1 = read only.  *old
 * (may be NULL) receives the previous value.  No reference counterpart
 * (a tuning knob; the autotuner searches it). */
int bsim_set_merge_width(int bits, int* old);

/* Kernel timing of the fused path (no reference counterpart; measurement
 * only): when enabled, every tile-pass launch is bracketed by HIP events on
 * the state's stream.  bsim_kernel_stats returns the launches and their
 * summed duration since the last reset. */
int bsim_set_profiling(int enable);
int bsim_kernel_stats(int reset, int64_t* launches, double* total_ms);

/* HBM stream probe (no reference counterpart; measurement only): allocates
 * a buffer of `bytes` (two for a copy), runs a plain streaming kernel over
 * it -- mode 0 copy (read + write every 16 B element), 1 read (reduce), 2
 * write (fill) -- `reps` times after one warm-up on `blocks` workgroups of
 * 256 threads (0 = 8 per CU), each thread moving 4, 8 or 16 x 16 B per
 * trip, and returns the fastest launch in *best_ms.  The ceiling the fused
 * passes and per-gate kernels are judged against.  bytes: a positive
 * multiple of 64 KiB. */
int bsim_stream_probe(int mode, int64_t bytes, int blocks, int reps, double* best_ms);

/* ---- noise (src/noise_kernels.jl) -------------------------------- */

/* apply_noise!(::Kraus, dm, targets...): src/noise_kernels.jl:105-196.
 * mode 0 = reference semantics (only blocks whose row and column agree on
 * every non-target bit; 2-qubit block order (00, t2, t1, 11), t1 = MSB;
 * k > 2: sorted endian targets), mode 1 = physical channel on every block
 * of the stored matrix (S <- sum conj(K) S K^T, S = rho^T).  k <= 8 in
 * mode 0, k <= 5 in mode 1. */
int bsim_dm_kraus(bsim_state* h, int k, const int* targets, int n_kraus, const double* kraus, int mode);
/* apply_noise!(::PhaseFlip, dm, q): src/noise_kernels.jl:7-23. */
int bsim_dm_phaseflip(bsim_state* h, int q, double p);
/* One superoperator pass: the fused form of a DM gate's two passes
 * (apply_gate!(U) on the column bits, apply_gate!(conj U) on the row bits,
 * src/dm_simulator.jl:130-140) and of a physical Kraus channel.  S is the
 * 4^k x 4^k column-major Liouville matrix of the map on the physical rho
 * under column stacking: rho'[r, c] = sum S[r + 2^k c, r' + 2^k c'] rho[r', c'],
 * matrix bit b of r and c <-> qubit targets[b] (so U rho U^dagger has
 * S = conj(U) (x) U, Kraus K_m have S = sum conj(K_m) (x) K_m).  k <= 5. */
int bsim_dm_apply_superop(bsim_state* h, int k, const int* targets, const double* S);

/* ---- results (src/result_types.jl, src/sv_simulator.jl:240-271,
 *      src/dm_simulator.jl:253-262) ------------------------------------ */

/* probabilities(): |psi_i|^2 or real(diag(rho)); out has 2^n doubles. */
int bsim_probabilities(bsim_state* h, double* out);
/* marginal_probability (src/result_types.jl:96-113); out has 2^nt doubles.
 * As in the reference the target order is ignored: the output index holds
 * the target qubits in ascending order, the smallest qubit number being
 * the most significant bit. */
int bsim_marginal_probabilities(bsim_state* h, int nt, const int* targets, double* out);
/* partial_trace (src/sv_simulator.jl:272-316, src/dm_simulator.jl:280-323):
 * the reduced density matrix of nt <= 8 target qubits, out = 4^nt complex
 * (column-major 2^nt x 2^nt, interleaved).  Output index bit nt-1-b holds
 * targets[b] (the first target is the most significant).  SV: out[i, j] =
 * sum_r psi[x(i, r)] conj(psi[x(j, r)]); DM: out[i, j] = sum_r
 * stored[x(i, r), x(j, r)], the stored matrix traced as the reference does.
 * nt = 0 gives the norm (SV) / trace (DM) in out[0].  The reference's own
 * reindexing of unsorted targets is applied above this call. */
int bsim_partial_trace(bsim_state* h, int nt, const int* targets, double* out);
/* Expectation of Pauli words (X/Y/Z/I tensor products, src/result_types.jl:
 * 161-167 with src/sv_simulator.jl:240-248, src/dm_simulator.jl:253-260).
 * Word w acts with Pauli paulis[w*n + q] ('I','X','Y','Z') on qubit q.
 * out[w] = <P_w> (real part). */
int bsim_expect_pauli(bsim_state* h, int n_words, const char* paulis, double* out);
/* <Z_q> for every qubit q in one pass: out has n doubles. */
int bsim_expect_z_all(bsim_state* h, double* out);
/* Expectation of a Hermitian matrix on k targets (src/sv_simulator.jl:162-230,
 * src/dm_simulator.jl:190-224): the reference applies it to a copy (1q:
 * standard; 2q: t1 = MSB; k > 2: sorted endian targets) and takes
 * real(<psi|O psi>) or real(tr).  k <= 6. */
int bsim_expect_matrix(bsim_state* h, int k, const int* targets, const double* m, double* out);
/* (Variance of a Hermitian observable, src/result_types.jl:184-194,223-238,
 * is <O*O> - <O>^2: the caller passes O*O to bsim_expect_matrix, as the
 * reference builds HermitianObservable(O*O).) */

/* ---- sampling (src/result_types.jl:8-80) ------------------------- */

#define BSIM_ALIAS_PARALLEL 0   /* exact integer sweep, fully parallel */
#define BSIM_ALIAS_REFERENCE 1  /* make_alias_table! order, one device thread */

/* Build the alias table of probabilities(h) on the device.  wsum <= 0
 * means: compute sum(p) on the device. */
int bsim_alias_build(bsim_state* h, int mode, double wsum);
/* Download the table: ap (2^n doubles) and alias (2^n int32, 0-based). */
int bsim_alias_download(bsim_state* h, double* ap, int32_t* alias);
/* Draw shots with Philox-4x32-10 keyed by seed, counter = shot index:
 * index = top n bits of (x1:x0), u = (x3:x2 >> 11) * 2^-53,
 * shot = u < ap[index] ? index : alias[index].  out: 0-based indices. */
int bsim_sample(bsim_state* h, int64_t shots, uint64_t seed, int64_t* out);
/* Same lookup with caller-supplied draws (0-based index, uniform). */
int bsim_sample_from_draws(bsim_state* h, int64_t shots, const int64_t* index, const double* u, int64_t* out);
/* _formatted_measurements (src/BraketSimulator.jl:88-101): shot indices to
 * bits of the measured qubits (qubit 0 = MSB), out has shots*nm int8. */
int bsim_shots_to_bits(bsim_state* h, int64_t shots, const int64_t* shot_index, int nm,
                       const int* measured, int8_t* out);

/* ---- differentiation (new: AdjointGradient, src/results.jl:90-160) --- */

/* Pauli-sum observable H = sum_w coeff[w] * P_w (paulis as in
 * bsim_expect_pauli).  Forward evolve from |0>, lambda = H psi, backward
 * sweep un-applying U^dagger to psi and lambda with the <lambda|dU|psi>
 * products fused into the gate kernel.  grad has n_params doubles
 * (op.param_index accumulates).  *expval = <H>.
 * On a density matrix (noise ops allowed; parameterized ops must be
 * BSIM_OP_GATE) the forward map cannot be uncomputed: rho is checkpointed
 * every s ops (s = ceil(sqrt(n_ops)) unless set below), each segment is
 * recomputed from its checkpoint, and the operator Lambda (initially H) is
 * propagated backwards through the adjoint maps (U^dagger . U; Kraus ops
 * K^dagger in the same mode); d<H>/dtheta = 2 Re tr(Lambda dU rho U^dagger).
 * The handle keeps the final state in both cases. */
int bsim_adjoint_gradient(bsim_state* h, const bsim_op* ops, int n_ops, int n_words, const char* paulis,
                          const double* coeffs, int n_params, double* grad, double* expval);
/* Checkpoint spacing (in ops) of the density-matrix adjoint; 0 = ceil(sqrt(n_ops)).
 * Returns the previous value in *old (may be NULL).  No reference
 * counterpart (the reference has no gradients). */
int bsim_set_checkpoint_interval(int ops, int* old);
/* State-vector adjoint backward sweep: 1 (default) = tile passes over the
 * reversed U^dagger ops with psi and lambda both in LDS and the parameter
 * products reduced inside the pass; 0 = one kernel per op.  Both give the
 * same gradient up to summation order.  With 1, states of fewer than 5
 * qubits, and sweeps whose per-tile partial sums would exceed max(4 GiB,
 * the state's bytes), still run one kernel per op.  Returns the previous
 * value in *old (may be NULL).  No reference counterpart. */
int bsim_set_adjoint_fusion(int on, int* old);
/* <H> for a Pauli sum on the current state. */
int bsim_expect_pauli_sum(bsim_state* h, int n_words, const char* paulis, const double* coeffs, double* out);

/* ---- batched execution (src/BraketSimulator.jl:347-395) ---------- */

typedef struct bsim_task {
    int32_t qubit_count;
    int32_t n_ops;
    const bsim_op* ops;
    int64_t shots;
    uint64_t seed;
    double* expect_z;      /* out: qubit_count doubles */
    int64_t* samples;      /* out: shots indices (may be NULL if shots == 0) */
    /* Gradient task (n_words > 0; parameter sweeps, no reference
     * counterpart): the circuit's parametrized ops carry
     * param_index / dmatrix as for bsim_adjoint_gradient, H = sum_w
     * coeffs[w] * P_w (paulis as in bsim_expect_pauli), shots must be 0 and
     * expect_z / samples are not used.  Outputs: grad (n_params doubles)
     * and *expval = <H> (may be NULL). */
    int32_t n_words;
    int32_t n_params;
    const char* paulis;
    const double* coeffs;
    double* grad;
    double* expval;
} bsim_task;

/* Run many state-vector circuits -- and adjoint gradients of parametrized
 * circuits -- in one persistent launch with a device work queue.
 * latency_us (optional, n_tasks doubles) receives each task's completion
 * time from the launch start.  A gradient task gives the gradient of
 * bsim_adjoint_gradient on the same circuit (same forward plan, backward
 * plan and per-wave product sums), <H> up to summation order. */
int bsim_batch_run(const bsim_task* tasks, int n_tasks, double* latency_us);

/* ---- captured evolutions and tuning ------------------------------ */

/* A captured evolve! (the per-simulation loop of src/sv_simulator.jl:140-148
 * and src/dm_simulator.jl:170-179, recorded once and replayed): the fused
 * plan of ops is built and its device tables uploaded once, then the launches
 * -- preceded by bsim_state_reset when reset != 0 -- are captured from the
 * state's stream into one hipGraph.  bsim_graph_launch replays it on the
 * state's stream (asynchronously; bsim_synchronize waits).  Only plans made
 * entirely of tile passes are captured (dense ops on <= 2 vector bits,
 * diagonals on <= 4, with any controls): anything else returns
 * BSIM_ENOTSUP.  The graph keeps h and the tile configuration of its
 * creation; destroy it before the state. */
typedef struct bsim_graph bsim_graph;
int bsim_graph_create(bsim_state* h, const bsim_op* ops, int n_ops, int reset, bsim_graph** out);
int bsim_graph_launch(bsim_graph* g);
int bsim_graph_destroy(bsim_graph* g);

/* Tile autotuner (the reference has no tunables besides CHUNK_SIZE,
 * src/BraketSimulator.jl:85-86): times the fused evolution of ops on h's
 * state (reset to |0...0> before every run, left in an unspecified state
 * afterwards) in a coordinate search: tile_bits 10..min(13, vector qubits)
 * with 3 or 4 sweep register bits at the current contiguous low bits, then
 * low bits 2..min(6, tile_bits-2) at the best shape, then the fusion width
 * (bsim_set_fusion_limit: 0 = unlimited, 48, 96, 192, 384 FP64 operations
 * per amplitude and pass), then -- when the plan has merged dense blocks --
 * their width (bsim_set_merge_width: off, 4, 5, 6 bits; read it back with
 * bsim_set_merge_width(-1, &w)); each candidate is one warm-up run and the
 * best of two timed runs (one when a run takes over 20 ms).  It keeps the
 * fastest as the library configuration and returns it (outputs may be
 * NULL; *best_ms = 0 when no candidate fits and the configuration is
 * unchanged).  Candidates are planned with private configurations: other
 * threads keep planning with the configuration in force until the winner
 * is published. */
int bsim_autotune(bsim_state* h, const bsim_op* ops, int n_ops, int* tile_bits, int* low_bits, int* sweep_bits,
                  int* fusion_limit, double* best_ms);

/* Grid cap of the per-gate kernels (their grid-stride loops): at most
 * `blocks` workgroups per launch, 0 = the default (64 per compute unit).
 * *old (may be NULL) receives the previous value.  A tuning and testing
 * knob with no reference counterpart: small caps make small states take
 * the multi-block-per-workgroup paths that large states take. */
int bsim_set_max_blocks(int blocks, int* old);

#ifdef __cplusplus
}
#endif

#endif /* BSIM_H */
