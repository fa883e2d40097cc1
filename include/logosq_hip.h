/*
 * logosq_hip.h -- C ABI of the MI355X (gfx950) dense state-vector backend.
 *
 * This is the drop-in boundary under LogosQ's State / Gate / Circuit /
 * Observable / GradientMethod API.  The reference crate (logosq 0.2.5) has no
 * FFI of its own: its boundary is the Rust trait API, and every entry point
 * below names the reference item it replaces (file:line in the reference
 * tree).  A Rust `extern "C"` block that binds these symbols is the only
 * glue a LogosQ maintainer needs.
 *
 * Conventions (identical to the reference):
 *   - amplitudes are complex, interleaved (re, im); complex128 = two f64
 *     (`num_complex::Complex64`, src/prelude.rs:180), complex64 = two f32;
 *   - qubit q is bit (n-1-q) of the basis index, so qubit 0 is the MSB
 *     (src/gates/utils.rs:5-7);
 *   - matrices are row-major, interleaved complex doubles.
 *
 * Errors: every function returns an lq_status and never aborts; the text of
 * the last error on the calling thread is available from lq_last_error().
 * Threading: calls on distinct handles may come from different threads.
 * All device work runs on one library stream with shared scratch buffers,
 * plan buffers and a shared cache of run-time compiled kernels, so every
 * entry point takes one library-wide lock and concurrent calls run one at a
 * time (recursive: an optimiser's objective callback may call back into the
 * library on the same thread).  One handle must not be used by two threads
 * at once (the reference's State is used through &mut).  Error text is kept
 * per thread.
 */
#ifndef LOGOSQ_HIP_H
#define LOGOSQ_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LQ_ABI_VERSION 1

/* Status codes mirror LogosQError (src/error.rs:9-62). */
typedef enum lq_status {
    LQ_OK = 0,
    LQ_ERR_INVALID_QUBIT_INDEX = 1,      /* InvalidQubitIndex      src/error.rs:11-12 */
    LQ_ERR_INVALID_STATE_DIMENSION = 2,  /* InvalidStateDimension  src/error.rs:15-16 */
    LQ_ERR_DIMENSION_MISMATCH = 3,       /* DimensionMismatch      src/error.rs:19-20 */
    LQ_ERR_INVALID_GATE_MATRIX = 4,      /* InvalidGateMatrix      src/error.rs:23-27 */
    LQ_ERR_PARAMETER_COUNT_MISMATCH = 5, /* ParameterCountMismatch src/error.rs:30-31 */
    LQ_ERR_CIRCUIT_QUBIT_MISMATCH = 6,   /* CircuitQubitMismatch   src/error.rs:34-40 */
    LQ_ERR_INVALID_ARGUMENT = 20,
    LQ_ERR_OUT_OF_MEMORY = 21,
    LQ_ERR_DEVICE = 22,
    LQ_ERR_UNSUPPORTED = 23
} lq_status;

typedef enum lq_dtype {
    LQ_C128 = 0, /* complex128: the reference's Complex64 */
    LQ_C64 = 1   /* complex64 (two f32); reductions still accumulate in f64 */
} lq_dtype;

/* Gate kinds.  Each names the reference struct whose `Gate::apply` loop it
 * replaces. */
typedef enum lq_gate_kind {
    LQ_GATE_X = 1,        /* XGate        src/gates/single_qubit.rs:14-31 */
    LQ_GATE_Y = 2,        /* YGate        src/gates/single_qubit.rs:39-59 (applies -Y) */
    LQ_GATE_Z = 3,        /* ZGate        src/gates/single_qubit.rs:67-82 */
    LQ_GATE_H = 4,        /* HGate        src/gates/single_qubit.rs:90-113 */
    LQ_GATE_S = 5,        /* SGate        src/gates/single_qubit.rs:121-137 */
    LQ_GATE_T = 6,        /* TGate        src/gates/single_qubit.rs:146-161 */
    LQ_GATE_RX = 10,      /* RXGate       src/gates/parameterized.rs:48-78 */
    LQ_GATE_RY = 11,      /* RYGate       src/gates/parameterized.rs:101-131 */
    LQ_GATE_RZ = 12,      /* RZGate       src/gates/parameterized.rs:154-176 */
    LQ_GATE_PHASE = 13,   /* PhaseGate    src/gates/parameterized.rs:199-218 */
    LQ_GATE_U3 = 14,      /* U3Gate       src/gates/parameterized.rs:249-282 */
    LQ_GATE_CNOT = 20,    /* CNOTGate     src/gates/two_qubit.rs:12-31 */
    LQ_GATE_SWAP = 21,    /* SWAPGate     src/gates/two_qubit.rs:40-62 */
    LQ_GATE_CZ = 22,      /* CZGate       src/gates/two_qubit.rs:71-90 */
    LQ_GATE_CPHASE = 23,  /* CPhaseGate   src/gates/two_qubit.rs:100-136 */
    LQ_GATE_CRX = 30,     /* CRXGate      src/gates/parameterized.rs:311-343 */
    LQ_GATE_CRY = 31,     /* CRYGate      src/gates/parameterized.rs:368-399 */
    LQ_GATE_CRZ = 32,     /* CRZGate      src/gates/parameterized.rs:424-451 */
    LQ_GATE_CP = 33,      /* CPhaseParamGate src/gates/parameterized.rs:476-499 */
    LQ_GATE_RXX = 40,     /* RXXGate      src/gates/parameterized.rs:528-578 */
    LQ_GATE_RYY = 41,     /* RYYGate      src/gates/parameterized.rs:604-653 */
    LQ_GATE_RZZ = 42,     /* RZZGate      src/gates/parameterized.rs:679-709 */
    LQ_GATE_TOFFOLI = 50, /* ToffoliGate  src/gates/three_qubit.rs:13-35 */
    /* Dense builder gates: Circuit::x/y/z/h/s/t/rx_matrix/... build a full
     * 2^n x 2^n MatrixGate (src/circuits/single_qubit.rs:12-63) whose apply is
     * `dot` followed by State::normalize (src/gates/mod.rs:77-93).  The
     * backend applies the 2x2 block to qubits[0] and renormalises, which is
     * the same operator without the O(4^n) matrix. */
    LQ_GATE_MATRIX1 = 60,
    /* Circuit::add_two_qubit_gate (src/circuits/two_qubits.rs:11-82): 4x4
     * block on the sorted pair (q1 < q2), sub-index (bit(q1) << 1) | bit(q2),
     * then renormalise. */
    LQ_GATE_MATRIX2 = 61
} lq_gate_kind;

/* Flag bits of lq_gate.flags. */
#define LQ_GATE_FLAG_RENORMALIZE 1 /* renormalise after the gate (MatrixGate semantics) */

/* One gate.  `theta[k]` is the k-th angle (theta, phi, lambda for U3).  When
 * `param[k] >= 0` the angle is parameter slot param[k] of an ansatz template
 * (ParameterizedCircuit, src/optimization/ansatz.rs:298-335) and theta[k] is
 * ignored by template-consuming entry points. */
typedef struct lq_gate {
    int32_t kind;
    int32_t qubits[3];
    double theta[3];
    int32_t param[3];
    int32_t flags;
    double matrix[32]; /* MATRIX1: 2x2 (8 doubles); MATRIX2: 4x4 (32 doubles) */
} lq_gate;

/* Pauli codes of Pauli (src/optimization/observable.rs:22-27). */
#define LQ_PAULI_I 0
#define LQ_PAULI_X 1
#define LQ_PAULI_Y 2
#define LQ_PAULI_Z 3

/* Y-phase convention of PauliTerm::expectation.  The reference takes the Y
 * phase from the row bit (src/optimization/observable.rs:125-133), which
 * returns (-1)^{#Y} <P>; LQ_PAULI_REFERENCE reproduces it bit for bit,
 * LQ_PAULI_PHYSICAL returns <P>. */
#define LQ_PAULI_REFERENCE 0
#define LQ_PAULI_PHYSICAL 1

/* Execution modes of lq_execute. */
#define LQ_EXEC_PER_GATE 0 /* one kernel per gate, in order (Gate::apply)          */
#define LQ_EXEC_FUSED 1    /* LDS tile passes that apply runs of gates per pass      */

typedef struct lq_state lq_state; /* opaque device-resident State */

/* ---- library ---------------------------------------------------------- */
const char* lq_last_error(void);
int lq_abi_version(void);
/* Device name, CU count, LDS bytes per block, global memory bytes. */
int lq_device_info(char* name, size_t name_len, int* cu_count, size_t* lds_bytes, size_t* global_bytes);
/* Block until all work queued by the library on its stream is done. */
int lq_synchronize(void);

/* Timing on the library stream (no counterpart in the reference): record a
 * HIP event in slot [0, 64) and read the time between two recorded slots. */
int lq_event_record(int slot);
int lq_event_elapsed_ms(int slot_start, int slot_end, float* ms);

/* ---- State (src/states/mod.rs:38-392) ---------------------------------- */
/* State::zero_state (mod.rs:136-141): |0...0>.  n in [1, 34]. */
int lq_state_create(int num_qubits, int dtype, lq_state** out);
int lq_state_destroy(lq_state* st);
int lq_state_num_qubits(const lq_state* st, int* num_qubits);
int lq_state_dtype(const lq_state* st, int* dtype);
/* zero_state / one_state (mod.rs:136-149): basis state |index>. */
int lq_state_set_basis(lq_state* st, uint64_t index);
/* plus_state (mod.rs:152-157): every amplitude 1/sqrt(2^n). */
int lq_state_set_plus(lq_state* st);
/* State::new / new_unchecked (mod.rs:55-86) when normalize != 0; a raw
 * vector_mut() write (mod.rs:131-133) when normalize == 0.  `amps` holds len
 * interleaved complex doubles; len must equal 2^n (InvalidStateDimension). */
int lq_state_upload(lq_state* st, const double* amps, uint64_t len, int normalize);
/* State::vector (mod.rs:122-124): copy the amplitudes to the host. */
int lq_state_download(const lq_state* st, double* amps, uint64_t len);
/* State: Clone (mod.rs:37).  dst and src must have the same n and dtype. */
int lq_state_copy(lq_state* dst, const lq_state* src);
/* State::normalize (mod.rs:91-114): divide by the norm when it exceeds 1e-10.
 * norm_out (nullable) receives the norm before scaling. */
int lq_state_normalize(lq_state* st, double* norm_out);
/* sum |a_i|^2 in f64. */
int lq_state_norm_sqr(const lq_state* st, double* out);
/* inner_product / inner_product_parallel (mod.rs:350-391): <a|b>. */
int lq_state_inner_product(const lq_state* a, const lq_state* b, double out[2]);
/* probabilities (mod.rs:277-279): len = 2^n doubles. */
int lq_state_probabilities(const lq_state* st, double* out, uint64_t len);

/* ---- Measurement (src/states/mod.rs:180-262, src/circuits/mod.rs:215-247)
 * The reference draws from thread_rng; here the caller supplies the uniforms
 * or a seed, so results are reproducible. */
/* p1[b] = probability that bit b of the basis index is 1, b = 0..n-1 (f64). */
int lq_bit_probabilities(const lq_state* st, double* p1);
/* State::measure_qubit (mod.rs:196-231): mask = 1 << qubit_idx (bit
 * qubit_idx of the index, as the reference does), outcome 1 when u < P(1),
 * then the projection and State::normalize.  u is a uniform in [0, 1). */
int lq_measure_qubit(lq_state* st, int qubit_idx, double u, int* result);
/* State::measure repeated `shots` times (mod.rs:180-185, 241-255): indices
 * drawn from |a_i|^2 (WeightedIndex rule: first i whose cumulative weight
 * exceeds u * total) with uniforms from SplitMix64(seed); no collapse. */
int lq_sample_shots(const lq_state* st, uint64_t seed, int64_t shots, uint64_t* out);

/* ---- Gates (Gate::apply, src/gates/mod.rs:26-35) ------------------------ */
/* Apply one gate with the kernel for its kind (the per-gate drop-in). */
int lq_apply_gate(lq_state* st, const lq_gate* gate);
/* QuantumStateBackend::apply_single_qubit_matrix without the renormalise
 * (src/states/backend.rs:23-24): 2x2 matrix m on `qubit`. */
int lq_apply_1q(lq_state* st, int qubit, const double m[8]);
/* 2x2 matrix on `target` in the subspace where every control qubit is 1. */
int lq_apply_ctrl_1q(lq_state* st, const int* controls, int num_controls, int target, const double m[8]);
/* QuantumStateBackend::apply_two_qubit_matrix without the renormalise
 * (src/states/mod.rs:441-456): 4x4 matrix, sub-index (bit(q_a) << 1) | bit(q_b). */
int lq_apply_2q(lq_state* st, int q_a, int q_b, const double m[32]);
/* MatrixGate::apply (src/gates/mod.rs:77-93): dense dim x dim matvec, then
 * State::normalize.  dim must equal 2^n (else DimensionMismatch). */
int lq_apply_dense(lq_state* st, const double* matrix, uint64_t dim);

/* ---- Circuits (src/circuits/mod.rs:352-375) ----------------------------- */
/* Circuit::execute: apply gates[0..count) in order.  mode is LQ_EXEC_*.
 * Angles come from theta[] (param[] is ignored). */
int lq_execute(lq_state* st, const lq_gate* gates, int count, int mode);

/* Tile geometry of the fused engine (no counterpart in the reference): tiles
 * of 2^tile_bits amplitudes per dtype whose low_bits lowest bits are always
 * contiguous; 0 keeps the default (12 / 13 / 5).  Used by the autotuner. */
int lq_set_fused_config(int tile_bits_c128, int tile_bits_c64, int low_bits);
/* Per-gate kernel variant (no counterpart in the reference): work units per
 * wavefront (1, 2, 4 or 8; default 1) and nontemporal global access on/off
 * (default on). */
int lq_set_gate_kernel_config(int unroll, int nontemporal);
/* Work-unit order of the per-gate kernels (no counterpart in the reference):
 * a mask of kernel classes whose work units run in a 32 x 32-transposed
 * order inside each block of 1024 units, so that 32 consecutive wavefronts
 * take units 32 apart (launches of at least 1024 units, i.e. 2^16 free
 * amplitudes; smaller ones keep the plain order).  Default LQ_UNITS_ALL; -1
 * restores the default.  Each unit is visited once in either order, so
 * the result is the same bit for bit; only the rate changes. */
#define LQ_UNITS_PAIR_HI 1  /* pair update, target above the lane bits      */
#define LQ_UNITS_PAIR_LO 2  /* pair update, target among the lane bits      */
#define LQ_UNITS_DIAG 4     /* diagonal gates                               */
#define LQ_UNITS_SWAP 8     /* SWAP                                         */
#define LQ_UNITS_QUAD 16    /* 4x4 blocks, both bits above the lane bits    */
#define LQ_UNITS_QUAD_LO 32 /* 4x4 blocks with a bit among the lane bits    */
#define LQ_UNITS_ALL 63
int lq_set_gate_unit_order(int classes);
/* Autotuner (no counterpart in the reference): times the per-gate kernel
 * variants and the fused tile geometries on a num_qubits state of dtype
 * (allocated for the call), keeps the fastest of each, and writes a JSON
 * report of every candidate into report (nullable, report_len bytes). */
int lq_autotune(int num_qubits, int dtype, char* report, size_t report_len);
/* Planner introspection (no counterpart in the reference): the order in which
 * the fused engine applies the gates (gate_order[k] = index of the k-th gate
 * applied, count entries), the pass each lands in, and the number of passes.
 * tile_bits / low_bits <= 0 select the defaults for dtype. */
int lq_plan_fused(int num_qubits, int dtype, const lq_gate* gates, int count, int tile_bits, int low_bits,
                  int* gate_order, int* pass_index, int* num_passes);

/* ---- Noise channels (src/noise/mod.rs:10-380) ---------------------------
 * Circuit::execute applies every noise model after every operation
 * (src/circuits/mod.rs:359-371).  `qubits` are the reference's qubit numbers
 * used as it uses them in these loops: bit `1 << qubit` of the basis index
 * (mod.rs:129,197,343-376), not bit n-1-qubit as for gates.  Indices >= n
 * are rejected (InvalidQubitIndex); probabilities outside [0, 1] are
 * rejected (InvalidArgument). */
/* AmplitudeDampingNoise::apply (mod.rs:110-147): out = sum over the listed
 * qubits of (K0 + K1) psi with K0 = |0><0| + sqrt(1-gamma)|1><1| and
 * K1 = sqrt(gamma)|0><1|, then State::normalize.  An empty list gives the
 * zero vector, as the reference's does. */
int lq_noise_amplitude_damping(lq_state* st, const int* qubits, int count, double gamma);
/* PhaseDampingNoise::apply (mod.rs:173-206): the last listed qubit's |1>
 * amplitudes times sqrt(1-lambda) (the reference resets its output from
 * psi for every qubit), then State::normalize; empty list: zero vector. */
int lq_noise_phase_damping(lq_state* st, const int* qubits, int count, double lambda);
/* DepolarizingNoise::apply (mod.rs:63-86): per listed qubit, when a uniform
 * u < p, a second uniform picks X, Y (the reference's i / -i form) or Z
 * (floor(3u), capped at 2).  Uniforms are the top 53 bits of SplitMix64
 * outputs from *rng_state, which is advanced; *errors (nullable) gets the
 * number of errors applied. */
int lq_noise_depolarizing(lq_state* st, const int* qubits, int count, double p, uint64_t* rng_state,
                          int* errors);

/* ---- Observables (src/optimization/observable.rs) ----------------------- */
/* PauliObservable::expectation (observable.rs:228-231) =
 * sum_t coeff[t] * PauliTerm::expectation (observable.rs:95-147).
 * paulis is num_terms x n codes (LQ_PAULI_*), row t for term t, column q for
 * qubit q.  convention is LQ_PAULI_REFERENCE or LQ_PAULI_PHYSICAL. */
int lq_expect_pauli_sum(const lq_state* st, const uint8_t* paulis, const double* coeffs, int num_terms,
                        int convention, double* out);
/* lq_expect_pauli_sum without allocating (no counterpart in the reference):
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and result as lq_expect_pauli_sum, but the term tables
 * and reduction partials live in a static 4 MiB device array that each GPU
 * holds from the library's load to the end of the process, and no pinned
 * host memory is taken.  A sum whose tables and partials need more than
 * 4 MiB on this state is rejected with LQ_ERR_INVALID_ARGUMENT before any
 * device work (the XYZ ring on 30 qubits needs about 2.1 MiB with LDS tiles
 * off).  Every call on a GPU shares that GPU's array; nothing in it has to
 * start at zero (each call writes all it reads).  Calls are serialised by
 * the library lock and wait for their kernels, so they do not overlap
 * within a process; device work of two calls that did overlap on one GPU
 * would race on the array. */
int lq_expect_pauli_sum_noalloc(const lq_state* st, const uint8_t* paulis, const double* coeffs, int num_terms,
                                int convention, double* out);
/* How flip groups (terms sharing their X/Y bits) are evaluated (no
 * reference counterpart): LQ_EXPECT_TILES_AUTO (default) reads a state over
 * 64 MiB once per tile layout, evaluating every flip group whose bits lie in
 * the layout's 2^11 (complex128) / 2^12 (complex64) amplitude tiles from
 * LDS; smaller states, and LQ_EXPECT_TILES_OFF, read the state once per
 * group; LQ_EXPECT_TILES_ON uses tiles at every size from 8 qubits.  Also
 * used by the batched gradients.  Results agree to rounding. */
#define LQ_EXPECT_TILES_OFF 0
#define LQ_EXPECT_TILES_AUTO 1
#define LQ_EXPECT_TILES_ON 2
int lq_set_expect_tiles(int mode);
/* The reads of the state an expectation of this sum makes in the current
 * modes: tile layouts; reads for the flip groups outside the tiles (orbit
 * items: up to four linearly independent flip masks per read); diagonal
 * groups read on their own.  With run-time specialisation on (lq_set_jit) and tiles
 * wanted for the size, the layouts are those of the specialised kernels
 * (128 KiB tiles, the diagonal group inside the first layout), used once
 * they compile. */
int lq_expect_plan(int num_qubits, int dtype, const uint8_t* paulis, const double* coeffs, int num_terms,
                   int convention, int* tile_layouts, int* pair_groups, int* diag_groups);
/* Compile the specialised expectation kernels of this sum on a num_qubits
 * state ahead of time (no device needed; no reference counterpart), so that
 * a timed call does not compile.  kernels: layouts compiled (0 when the
 * modes do not use them at this size). */
int lq_expect_jit_compile(int num_qubits, int dtype, const uint8_t* paulis, const double* coeffs, int num_terms,
                          int convention, int* kernels);
/* The specialised plan of this sum as JSON (introspection, no reference
 * counterpart): per layout its tile geometry, coefficient table and kernel
 * text, and the flip groups left to the pair kernel; "{}" when the modes do
 * not use it at this size.  needed (nullable) gets the buffer size. */
int lq_expect_jit_describe(int num_qubits, int dtype, const uint8_t* paulis, const double* coeffs, int num_terms,
                           int convention, char* out, size_t out_len, size_t* needed);

/* ---- Gradients (src/optimization/gradient.rs) --------------------------- */
/* ParameterShift::compute_gradient (gradient.rs:45-78) for an ansatz given as
 * a gate template (param[] slots) of num_params parameters on n qubits,
 * evaluated from |0...0> for the Pauli-sum observable.  energy_out (nullable)
 * receives E(params).  mode: LQ_GRAD_SEQUENTIAL runs the 2P circuits one
 * after another with per-gate kernels, as the reference does;
 * LQ_GRAD_BATCHED runs all 2P+1 circuits together in tile passes. */
#define LQ_GRAD_SEQUENTIAL 0
#define LQ_GRAD_BATCHED 1
/* grad_mode of lq_vqe_run only: the adjoint gradient below. */
#define LQ_GRAD_ADJOINT 2
int lq_param_shift_gradient(int num_qubits, int dtype, const lq_gate* tmpl, int num_gates, const double* params,
                            int num_params, double shift, const uint8_t* paulis, const double* coeffs,
                            int num_terms, int convention, int mode, double* energy_out, double* grad_out);

/* Adjoint-method gradient (no counterpart in the reference, which offers
 * ParameterShift and FiniteDifference): one forward sweep, lambda = H psi,
 * one backward sweep; dE/dtheta = Im <lambda|P|psi> at each rotation
 * exp(-i theta/2 P).  Parameters may sit on RX / RY / RZ only; every gate
 * must have a unitary inverse (no renormalising MatrixGate).  Equals
 * ParameterShift (shift pi/2) up to rounding. */
int lq_adjoint_gradient(int num_qubits, int dtype, const lq_gate* tmpl, int num_gates, const double* params,
                        int num_params, const uint8_t* paulis, const double* coeffs, int num_terms, int convention,
                        double* energy_out, double* grad_out);
/* The adjoint gradient of `count` instances of one template at once (no
 * reference counterpart): instance i has parameters params[i*num_params ..]
 * and Hamiltonian coefficients coeffs[i*num_terms ..] over the shared Pauli
 * strings `paulis`.  energies (nullable) gets count values, grads
 * count x num_params.  One launch per tile pass serves all instances. */
int lq_adjoint_gradient_batch(int num_qubits, int dtype, const lq_gate* tmpl, int num_gates, int count,
                              const double* params, int num_params, const uint8_t* paulis, const double* coeffs,
                              int num_terms, int convention, double* energies, double* grads);

/* ---- Batches (no reference counterpart: the reference evaluates circuits
 * one after another) ------------------------------------------------------ */
/* count circuits, circuit c on num_qubits[c] (<= 30) qubits with gates
 * circuits[c][0 .. num_gates[c]) (angles from theta[]), each run from
 * |0...0> in dtype and measured with its own Pauli sum (paulis[c]: a
 * num_terms[c] x num_qubits[c] code matrix, coeffs[c]).  Circuits of one
 * structure share one pass plan and run together, one launch per pass;
 * every expectation is reduced in f64.  energies[c] receives the result. */
int lq_batch_expectations(int count, const int* num_qubits, const lq_gate* const* circuits, const int* num_gates,
                          const uint8_t* const* paulis, const double* const* coeffs, const int* num_terms,
                          int convention, int dtype, double* energies);

/* Batches (parameter-shift gradients and lq_batch_expectations) run their
 * device work as one launch sequence per chunk; with graphs on (default) a
 * sequence shape seen a second time is captured into a hipGraph and replayed
 * from then on (no reference counterpart).  Results are identical either way. */
int lq_set_batch_graphs(int enable);
/* Graph captures and replays so far in this process. */
int lq_batch_graph_stats(long long* captures, long long* replays);

/* Run-time specialisation of fused tile passes (no reference counterpart).
 * A pass kernel is generated for each pass structure (tile geometry, gate
 * kinds, positions and control masks; not the angles), compiled with hipRTC
 * and launched in place of the generic tile kernel.  LQ_JIT_REPEAT (the
 * default) compiles a structure the second time a plan holds it, for passes
 * of at least 1024 tiles; LQ_JIT_ALWAYS compiles every structure at first
 * sight; LQ_JIT_OFF keeps the generic kernel.  The arithmetic is the same,
 * so results do not depend on the mode. */
#define LQ_JIT_OFF 0
#define LQ_JIT_REPEAT 1
#define LQ_JIT_ALWAYS 2
int lq_set_jit(int mode);
/* Kernels compiled, compiles or loads that failed (those passes stay on the
 * generic kernel), specialised launches, and seconds spent compiling. */
int lq_jit_stats(long long* compiled, long long* failed, long long* launches, double* compile_seconds);
/* Compile the specialised kernels of a circuit's fused plan ahead of time
 * (needs no device; loaded on first use).  kernels: kernels ready. */
int lq_jit_compile_plan(int num_qubits, int dtype, const lq_gate* gates, int count, int* kernels);

/* ---- Optimizers and VQE (src/optimization/optimizer.rs, vqe.rs) -------- */
typedef double (*lq_objective_fn)(const double* x, int dim, void* user);
typedef void (*lq_gradient_fn)(const double* x, int dim, double* grad, void* user);
/* OptimizerResult (src/optimization/optimizer.rs:8-15); the parameters are
 * returned in place and the convergence history in a caller buffer. */
typedef struct lq_opt_result {
    double optimal_value;
    int32_t num_iterations; /* length of the convergence history */
    int32_t converged;
} lq_opt_result;
/* GradientDescent::minimize (optimizer.rs:56-107).  x: initial point in,
 * optimal parameters out; history (nullable) holds max_iterations doubles. */
int lq_gd_minimize(double learning_rate, int max_iterations, double tolerance, lq_objective_fn f, lq_gradient_fn g,
                   void* user, double* x, int dim, double* history, lq_opt_result* result);
/* Adam::minimize (optimizer.rs:139-223). */
int lq_adam_minimize(double learning_rate, double beta1, double beta2, double epsilon, int max_iterations,
                     double tolerance, lq_objective_fn f, lq_gradient_fn g, void* user, double* x, int dim,
                     double* history, lq_opt_result* result);
/* VQE::run (src/optimization/vqe.rs:55-99) with ParameterShift gradients
 * (grad_mode LQ_GRAD_*) and GradientDescent (optimizer 0) or Adam
 * (optimizer 1, beta1 0.9, beta2 0.999, eps 1e-8).  params: initial in,
 * optimal out.  final_state (nullable) receives the state at the optimum. */
int lq_vqe_run(int num_qubits, int dtype, const lq_gate* tmpl, int num_gates, int num_params,
               const uint8_t* paulis, const double* coeffs, int num_terms, int convention, double shift,
               int grad_mode, int optimizer, double learning_rate, int max_iterations, double tolerance,
               double* params, double* history, lq_state* final_state, lq_opt_result* result);
/* VQE::run with Adam (grad_mode LQ_GRAD_*) that also stops as soon as a
 * recorded energy is at or below target_energy (no reference counterpart:
 * time-to-target measurements).  Until then every iteration is exactly
 * VQE::run's; *reached (nullable) is 1 when the target stopped the run.
 * The history holds the energies up to and including the first one at or
 * below the target. */
int lq_vqe_run_to_target(int num_qubits, int dtype, const lq_gate* tmpl, int num_gates, int num_params,
                         const uint8_t* paulis, const double* coeffs, int num_terms, int convention, double shift,
                         int grad_mode, double learning_rate, int max_iterations, double tolerance,
                         double target_energy, double* params, double* history, lq_opt_result* result,
                         int32_t* reached);
/* `count` VQE::run calls with Adam and the adjoint gradient, advanced in
 * lock step (no reference counterpart): instance i starts from
 * params[i*num_params ..] and minimises the Hamiltonian with coefficients
 * coeffs[i*num_terms ..] over the shared Pauli strings.  Each instance
 * follows Adam::minimize (optimizer.rs:139-223) exactly, with its own
 * stopping tests; every iteration evaluates all unfinished instances with
 * one lq_adjoint_gradient_batch call.  Outputs per instance: optimal
 * parameters (in place), history (nullable, count x max_iterations),
 * energies (optimal value), iterations (history length), converged. */
int lq_vqe_adam_batch(int num_qubits, int dtype, const lq_gate* tmpl, int num_gates, int num_params,
                      const uint8_t* paulis, const double* coeffs, int num_terms, int convention, int count,
                      double learning_rate, int max_iterations, double tolerance, double* params, double* history,
                      double* energies, int32_t* iterations, int32_t* converged);

/* ---- Warm-start network (no reference counterpart: the reference starts
 * every VQE from caller-given parameters, src/optimization/vqe.rs:55-60) --
 * A perceptron in -> hidden -> hidden -> out with GELU (erf form) after the
 * two hidden layers.  f32 master weights; every matrix product rounds its
 * operands to bf16 and accumulates in f32 (MFMA).  Parameters are one f32
 * vector: W1[hidden][in], b1, W2[hidden][hidden], b2, W3[out][hidden], b3.
 * Training minimises the mean squared error with Adam (beta1 0.9, beta2
 * 0.999, eps 1e-8) on the device. */
typedef struct lq_mlp lq_mlp;
/* in_dim <= 8, 16 <= hidden <= 4096, out_dim <= 8192; weights and biases
 * uniform in +-1/sqrt(fan_in) from SplitMix64(seed). */
int lq_mlp_create(int in_dim, int hidden, int out_dim, uint64_t seed, lq_mlp** out);
int lq_mlp_destroy(lq_mlp* net);
int lq_mlp_num_params(const lq_mlp* net, int64_t* count);
int lq_mlp_get_params(const lq_mlp* net, float* out, int64_t count);
/* Sets the weights and resets the Adam moments. */
int lq_mlp_set_params(lq_mlp* net, const float* in, int64_t count);
/* y[rows][out] = net(x[rows][in]) */
int lq_mlp_forward(lq_mlp* net, const float* x, int rows, float* y);
/* Mean squared error over rows x out and its gradient (no update);
 * loss and grad (num_params floats) are nullable. */
int lq_mlp_gradients(lq_mlp* net, const float* x, const float* target, int rows, float* loss, float* grad);
/* epochs passes of Adam minibatch steps over count samples (shuffled per
 * epoch by SplitMix64(seed)); loss_history (nullable) receives the mean
 * minibatch loss of each epoch. */
int lq_mlp_train(lq_mlp* net, const float* x, const float* target, int count, int batch, int epochs,
                 float learning_rate, uint64_t seed, float* loss_history);

#ifdef __cplusplus
}
#endif

#endif /* LOGOSQ_HIP_H */
