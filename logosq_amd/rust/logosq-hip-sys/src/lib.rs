//! Raw bindings to `liblogosq_hip.so`, the MI355X (gfx950) state-vector
//! backend, one declaration per entry point of `include/logosq_hip.h` (same
//! names, argument order and meaning; every function returns an `lq_status`
//! code and `lq_last_error` gives the message of the calling thread's last
//! failure).  The entry points replace the hot loops of the logosq crate:
//! `Gate::apply` (src/gates/mod.rs:26-35), `Circuit::execute`
//! (src/circuits/mod.rs:352-375), `PauliObservable::expectation`
//! (src/optimization/observable.rs:95-147, 228-231), `ParameterShift`
//! (src/optimization/gradient.rs:31-73) and `VQE::run`
//! (src/optimization/vqe.rs:55-99).
#![allow(non_camel_case_types)]

use std::ffi::CStr;
use std::os::raw::{c_char, c_int, c_longlong, c_void};

pub const LQ_ABI_VERSION: c_int = 1;

// lq_status, mirroring LogosQError (src/error.rs:9-62)
pub const LQ_OK: c_int = 0;
pub const LQ_ERR_INVALID_QUBIT_INDEX: c_int = 1;
pub const LQ_ERR_INVALID_STATE_DIMENSION: c_int = 2;
pub const LQ_ERR_DIMENSION_MISMATCH: c_int = 3;
pub const LQ_ERR_INVALID_GATE_MATRIX: c_int = 4;
pub const LQ_ERR_PARAMETER_COUNT_MISMATCH: c_int = 5;
pub const LQ_ERR_CIRCUIT_QUBIT_MISMATCH: c_int = 6;
pub const LQ_ERR_INVALID_ARGUMENT: c_int = 20;
pub const LQ_ERR_OUT_OF_MEMORY: c_int = 21;
pub const LQ_ERR_DEVICE: c_int = 22;
pub const LQ_ERR_UNSUPPORTED: c_int = 23;

// lq_dtype
pub const LQ_C128: c_int = 0;
pub const LQ_C64: c_int = 1;

// lq_gate_kind: the gate structs of src/gates/*.rs
pub const LQ_GATE_X: i32 = 1;
pub const LQ_GATE_Y: i32 = 2;
pub const LQ_GATE_Z: i32 = 3;
pub const LQ_GATE_H: i32 = 4;
pub const LQ_GATE_S: i32 = 5;
pub const LQ_GATE_T: i32 = 6;
pub const LQ_GATE_RX: i32 = 10;
pub const LQ_GATE_RY: i32 = 11;
pub const LQ_GATE_RZ: i32 = 12;
pub const LQ_GATE_PHASE: i32 = 13;
pub const LQ_GATE_U3: i32 = 14;
pub const LQ_GATE_CNOT: i32 = 20;
pub const LQ_GATE_SWAP: i32 = 21;
pub const LQ_GATE_CZ: i32 = 22;
pub const LQ_GATE_CPHASE: i32 = 23;
pub const LQ_GATE_CRX: i32 = 30;
pub const LQ_GATE_CRY: i32 = 31;
pub const LQ_GATE_CRZ: i32 = 32;
pub const LQ_GATE_CP: i32 = 33;
pub const LQ_GATE_RXX: i32 = 40;
pub const LQ_GATE_RYY: i32 = 41;
pub const LQ_GATE_RZZ: i32 = 42;
pub const LQ_GATE_TOFFOLI: i32 = 50;
pub const LQ_GATE_MATRIX1: i32 = 60;
pub const LQ_GATE_MATRIX2: i32 = 61;
pub const LQ_GATE_FLAG_RENORMALIZE: i32 = 1;

// Pauli codes (src/optimization/observable.rs:22-27) and Y conventions
pub const LQ_PAULI_I: u8 = 0;
pub const LQ_PAULI_X: u8 = 1;
pub const LQ_PAULI_Y: u8 = 2;
pub const LQ_PAULI_Z: u8 = 3;
pub const LQ_PAULI_REFERENCE: c_int = 0;
pub const LQ_PAULI_PHYSICAL: c_int = 1;

pub const LQ_EXEC_PER_GATE: c_int = 0;
pub const LQ_EXEC_FUSED: c_int = 1;
pub const LQ_EXPECT_TILES_OFF: c_int = 0;
pub const LQ_EXPECT_TILES_AUTO: c_int = 1;
pub const LQ_EXPECT_TILES_ON: c_int = 2;
pub const LQ_GRAD_SEQUENTIAL: c_int = 0;
pub const LQ_GRAD_BATCHED: c_int = 1;
pub const LQ_GRAD_ADJOINT: c_int = 2;
pub const LQ_UNITS_PAIR_HI: c_int = 1;
pub const LQ_UNITS_PAIR_LO: c_int = 2;
pub const LQ_UNITS_DIAG: c_int = 4;
pub const LQ_UNITS_SWAP: c_int = 8;
pub const LQ_UNITS_QUAD: c_int = 16;
pub const LQ_UNITS_QUAD_LO: c_int = 32;
pub const LQ_UNITS_ALL: c_int = 63;
pub const LQ_JIT_OFF: c_int = 0;
pub const LQ_JIT_REPEAT: c_int = 1;
pub const LQ_JIT_ALWAYS: c_int = 2;

/// lq_gate (312 bytes).  `param[k] >= 0` makes angle k the parameter slot
/// param[k] of an ansatz template (ParameterizedCircuit,
/// src/optimization/ansatz.rs:298-335); `matrix` is row-major interleaved
/// complex (2x2 for MATRIX1, 4x4 for MATRIX2).
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct LqGate {
    pub kind: i32,
    pub qubits: [i32; 3],
    pub theta: [f64; 3],
    pub param: [i32; 3],
    pub flags: i32,
    pub matrix: [f64; 32],
}

impl LqGate {
    /// A gate of `kind` on `qubits` (unused slots -1) with fixed angles.
    pub fn new(kind: i32, qubits: &[i32], theta: &[f64]) -> LqGate {
        let mut g = LqGate { kind, qubits: [-1; 3], theta: [0.0; 3], param: [-1; 3], flags: 0, matrix: [0.0; 32] };
        g.qubits[..qubits.len()].copy_from_slice(qubits);
        g.theta[..theta.len()].copy_from_slice(theta);
        g
    }
}

/// OptimizerResult (src/optimization/optimizer.rs:8-15).
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct LqOptResult {
    pub optimal_value: f64,
    pub num_iterations: i32,
    pub converged: i32,
}

/// Opaque device-resident State.
#[repr(C)]
pub struct LqState {
    _private: [u8; 0],
}

/// Opaque warm-start network.
#[repr(C)]
pub struct LqMlp {
    _private: [u8; 0],
}

pub type LqObjectiveFn = Option<unsafe extern "C" fn(x: *const f64, dim: c_int, user: *mut c_void) -> f64>;
pub type LqGradientFn = Option<unsafe extern "C" fn(x: *const f64, dim: c_int, grad: *mut f64, user: *mut c_void)>;

extern "C" {
    // ---- library
    pub fn lq_last_error() -> *const c_char;
    pub fn lq_abi_version() -> c_int;
    pub fn lq_device_info(name: *mut c_char, name_len: usize, cu_count: *mut c_int, lds_bytes: *mut usize, global_bytes: *mut usize) -> c_int;
    pub fn lq_synchronize() -> c_int;
    pub fn lq_event_record(slot: c_int) -> c_int;
    pub fn lq_event_elapsed_ms(slot_start: c_int, slot_end: c_int, ms: *mut f32) -> c_int;

    // ---- State (src/states/mod.rs:38-392)
    pub fn lq_state_create(num_qubits: c_int, dtype: c_int, out: *mut *mut LqState) -> c_int;
    pub fn lq_state_destroy(st: *mut LqState) -> c_int;
    pub fn lq_state_num_qubits(st: *const LqState, num_qubits: *mut c_int) -> c_int;
    pub fn lq_state_dtype(st: *const LqState, dtype: *mut c_int) -> c_int;
    pub fn lq_state_set_basis(st: *mut LqState, index: u64) -> c_int;
    pub fn lq_state_set_plus(st: *mut LqState) -> c_int;
    pub fn lq_state_upload(st: *mut LqState, amps: *const f64, len: u64, normalize: c_int) -> c_int;
    pub fn lq_state_download(st: *const LqState, amps: *mut f64, len: u64) -> c_int;
    pub fn lq_state_copy(dst: *mut LqState, src: *const LqState) -> c_int;
    pub fn lq_state_normalize(st: *mut LqState, norm_out: *mut f64) -> c_int;
    pub fn lq_state_norm_sqr(st: *const LqState, out: *mut f64) -> c_int;
    pub fn lq_state_inner_product(a: *const LqState, b: *const LqState, out: *mut f64) -> c_int;
    pub fn lq_state_probabilities(st: *const LqState, out: *mut f64, len: u64) -> c_int;

    // ---- Measurement (src/states/mod.rs:180-262, src/circuits/mod.rs:215-247)
    pub fn lq_bit_probabilities(st: *const LqState, p1: *mut f64) -> c_int;
    pub fn lq_measure_qubit(st: *mut LqState, qubit_idx: c_int, u: f64, result: *mut c_int) -> c_int;
    pub fn lq_sample_shots(st: *const LqState, seed: u64, shots: i64, out: *mut u64) -> c_int;

    // ---- Gates (Gate::apply, src/gates/mod.rs:26-35)
    pub fn lq_apply_gate(st: *mut LqState, gate: *const LqGate) -> c_int;
    pub fn lq_apply_1q(st: *mut LqState, qubit: c_int, m: *const f64) -> c_int;
    pub fn lq_apply_ctrl_1q(st: *mut LqState, controls: *const c_int, num_controls: c_int, target: c_int, m: *const f64) -> c_int;
    pub fn lq_apply_2q(st: *mut LqState, q_a: c_int, q_b: c_int, m: *const f64) -> c_int;
    pub fn lq_apply_dense(st: *mut LqState, matrix: *const f64, dim: u64) -> c_int;

    // ---- Circuits (src/circuits/mod.rs:352-375)
    pub fn lq_execute(st: *mut LqState, gates: *const LqGate, count: c_int, mode: c_int) -> c_int;
    pub fn lq_set_fused_config(tile_bits_c128: c_int, tile_bits_c64: c_int, low_bits: c_int) -> c_int;
    pub fn lq_set_gate_kernel_config(unroll: c_int, nontemporal: c_int) -> c_int;
    pub fn lq_set_gate_unit_order(classes: c_int) -> c_int;
    pub fn lq_autotune(num_qubits: c_int, dtype: c_int, report: *mut c_char, report_len: usize) -> c_int;
    pub fn lq_plan_fused(num_qubits: c_int, dtype: c_int, gates: *const LqGate, count: c_int, tile_bits: c_int, low_bits: c_int, gate_order: *mut c_int, pass_index: *mut c_int, num_passes: *mut c_int) -> c_int;

    // ---- Noise channels (src/noise/mod.rs:10-380)
    pub fn lq_noise_amplitude_damping(st: *mut LqState, qubits: *const c_int, count: c_int, gamma: f64) -> c_int;
    pub fn lq_noise_phase_damping(st: *mut LqState, qubits: *const c_int, count: c_int, lambda: f64) -> c_int;
    pub fn lq_noise_depolarizing(st: *mut LqState, qubits: *const c_int, count: c_int, p: f64, rng_state: *mut u64, errors: *mut c_int) -> c_int;

    // ---- Observables (src/optimization/observable.rs)
    pub fn lq_expect_pauli_sum(st: *const LqState, paulis: *const u8, coeffs: *const f64, num_terms: c_int, convention: c_int, out: *mut f64) -> c_int;
    pub fn lq_expect_pauli_sum_noalloc(st: *const LqState, paulis: *const u8, coeffs: *const f64, num_terms: c_int, convention: c_int, out: *mut f64) -> c_int;
    pub fn lq_set_expect_tiles(mode: c_int) -> c_int;
    pub fn lq_expect_plan(num_qubits: c_int, dtype: c_int, paulis: *const u8, coeffs: *const f64, num_terms: c_int, convention: c_int, tile_layouts: *mut c_int, pair_groups: *mut c_int, diag_groups: *mut c_int) -> c_int;
    pub fn lq_expect_jit_compile(num_qubits: c_int, dtype: c_int, paulis: *const u8, coeffs: *const f64, num_terms: c_int, convention: c_int, kernels: *mut c_int) -> c_int;
    pub fn lq_expect_jit_describe(num_qubits: c_int, dtype: c_int, paulis: *const u8, coeffs: *const f64, num_terms: c_int, convention: c_int, out: *mut c_char, out_len: usize, needed: *mut usize) -> c_int;

    // ---- Gradients (src/optimization/gradient.rs)
    pub fn lq_param_shift_gradient(num_qubits: c_int, dtype: c_int, tmpl: *const LqGate, num_gates: c_int, params: *const f64, num_params: c_int, shift: f64, paulis: *const u8, coeffs: *const f64, num_terms: c_int, convention: c_int, mode: c_int, energy_out: *mut f64, grad_out: *mut f64) -> c_int;
    pub fn lq_adjoint_gradient(num_qubits: c_int, dtype: c_int, tmpl: *const LqGate, num_gates: c_int, params: *const f64, num_params: c_int, paulis: *const u8, coeffs: *const f64, num_terms: c_int, convention: c_int, energy_out: *mut f64, grad_out: *mut f64) -> c_int;
    pub fn lq_adjoint_gradient_batch(num_qubits: c_int, dtype: c_int, tmpl: *const LqGate, num_gates: c_int, count: c_int, params: *const f64, num_params: c_int, paulis: *const u8, coeffs: *const f64, num_terms: c_int, convention: c_int, energies: *mut f64, grads: *mut f64) -> c_int;

    // ---- Batches (no reference counterpart)
    pub fn lq_batch_expectations(count: c_int, num_qubits: *const c_int, circuits: *const *const LqGate, num_gates: *const c_int, paulis: *const *const u8, coeffs: *const *const f64, num_terms: *const c_int, convention: c_int, dtype: c_int, energies: *mut f64) -> c_int;
    pub fn lq_set_batch_graphs(enable: c_int) -> c_int;
    pub fn lq_batch_graph_stats(captures: *mut c_longlong, replays: *mut c_longlong) -> c_int;
    pub fn lq_set_jit(mode: c_int) -> c_int;
    pub fn lq_jit_stats(compiled: *mut c_longlong, failed: *mut c_longlong, launches: *mut c_longlong, compile_seconds: *mut f64) -> c_int;
    pub fn lq_jit_compile_plan(num_qubits: c_int, dtype: c_int, gates: *const LqGate, count: c_int, kernels: *mut c_int) -> c_int;

    // ---- Optimizers and VQE (src/optimization/optimizer.rs, vqe.rs)
    pub fn lq_gd_minimize(learning_rate: f64, max_iterations: c_int, tolerance: f64, f: LqObjectiveFn, g: LqGradientFn, user: *mut c_void, x: *mut f64, dim: c_int, history: *mut f64, result: *mut LqOptResult) -> c_int;
    pub fn lq_adam_minimize(learning_rate: f64, beta1: f64, beta2: f64, epsilon: f64, max_iterations: c_int, tolerance: f64, f: LqObjectiveFn, g: LqGradientFn, user: *mut c_void, x: *mut f64, dim: c_int, history: *mut f64, result: *mut LqOptResult) -> c_int;
    pub fn lq_vqe_run(num_qubits: c_int, dtype: c_int, tmpl: *const LqGate, num_gates: c_int, num_params: c_int, paulis: *const u8, coeffs: *const f64, num_terms: c_int, convention: c_int, shift: f64, grad_mode: c_int, optimizer: c_int, learning_rate: f64, max_iterations: c_int, tolerance: f64, params: *mut f64, history: *mut f64, final_state: *mut LqState, result: *mut LqOptResult) -> c_int;
    pub fn lq_vqe_run_to_target(num_qubits: c_int, dtype: c_int, tmpl: *const LqGate, num_gates: c_int, num_params: c_int, paulis: *const u8, coeffs: *const f64, num_terms: c_int, convention: c_int, shift: f64, grad_mode: c_int, learning_rate: f64, max_iterations: c_int, tolerance: f64, target_energy: f64, params: *mut f64, history: *mut f64, result: *mut LqOptResult, reached: *mut i32) -> c_int;
    pub fn lq_vqe_adam_batch(num_qubits: c_int, dtype: c_int, tmpl: *const LqGate, num_gates: c_int, num_params: c_int, paulis: *const u8, coeffs: *const f64, num_terms: c_int, convention: c_int, count: c_int, learning_rate: f64, max_iterations: c_int, tolerance: f64, params: *mut f64, history: *mut f64, energies: *mut f64, iterations: *mut i32, converged: *mut i32) -> c_int;

    // ---- Warm-start network (no reference counterpart)
    pub fn lq_mlp_create(in_dim: c_int, hidden: c_int, out_dim: c_int, seed: u64, out: *mut *mut LqMlp) -> c_int;
    pub fn lq_mlp_destroy(net: *mut LqMlp) -> c_int;
    pub fn lq_mlp_num_params(net: *const LqMlp, count: *mut i64) -> c_int;
    pub fn lq_mlp_get_params(net: *const LqMlp, out: *mut f32, count: i64) -> c_int;
    pub fn lq_mlp_set_params(net: *mut LqMlp, in_: *const f32, count: i64) -> c_int;
    pub fn lq_mlp_forward(net: *mut LqMlp, x: *const f32, rows: c_int, y: *mut f32) -> c_int;
    pub fn lq_mlp_gradients(net: *mut LqMlp, x: *const f32, target: *const f32, rows: c_int, loss: *mut f32, grad: *mut f32) -> c_int;
    pub fn lq_mlp_train(net: *mut LqMlp, x: *const f32, target: *const f32, count: c_int, batch: c_int, epochs: c_int, learning_rate: f32, seed: u64, loss_history: *mut f32) -> c_int;
}

/// A failed call: its status code and the library's message.
#[derive(Debug, Clone)]
pub struct LqError {
    pub status: c_int,
    pub message: String,
}

/// Turns a status code into a Result, reading `lq_last_error` on failure.
pub fn check(status: c_int) -> Result<(), LqError> {
    if status == LQ_OK {
        return Ok(());
    }
    let message = unsafe {
        let p = lq_last_error();
        if p.is_null() { String::new() } else { CStr::from_ptr(p).to_string_lossy().into_owned() }
    };
    Err(LqError { status, message })
}

#[cfg(test)]
mod tests {
    use super::*;

    #[test]
    fn struct_layouts_match_the_header() {
        assert_eq!(std::mem::size_of::<LqGate>(), 312);
        assert_eq!(std::mem::align_of::<LqGate>(), 8);
        assert_eq!(std::mem::size_of::<LqOptResult>(), 16);
    }
}
