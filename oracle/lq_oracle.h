/*
 * lq_oracle.h -- CPU restatement of LogosQ 0.2.5's dense state-vector hot
 * path.  TEST INFRASTRUCTURE ONLY: it is the parity oracle and the CPU
 * baseline, never part of the product path.
 *
 * Every function follows one reference loop (file:line given per function)
 * with the same operation order and the same complex arithmetic as
 * num-complex 0.4: (a+bi)(c+di) = (ac - bd) + (ad + bc)i, each product and
 * sum rounded separately (built with -ffp-contract=off so no FMA is formed).
 * Amplitudes are interleaved doubles; qubit q is bit n-1-q.
 */
#ifndef LQ_ORACLE_H
#define LQ_ORACLE_H

#include <stdint.h>

#include "../include/logosq_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double re, im;
} orc;

/* State */
double or_normalize(orc* v, uint64_t dim);            /* src/states/mod.rs:91-114 */
void or_zero_state(orc* v, int n);                    /* src/states/mod.rs:136-141 */
void or_one_state(orc* v, int n);                     /* src/states/mod.rs:144-149 */
void or_plus_state(orc* v, int n);                    /* src/states/mod.rs:152-157 */
void or_inner_product(const orc* a, const orc* b, uint64_t dim, double out[2]); /* mod.rs:377-391 */

/* Gates: the reference's Gate::apply loops. */
void or_x(orc* v, int n, int q);                      /* single_qubit.rs:14-31 */
void or_y(orc* v, int n, int q);                      /* single_qubit.rs:39-59 */
void or_z(orc* v, int n, int q);                      /* single_qubit.rs:67-82 */
void or_h(orc* v, int n, int q);                      /* single_qubit.rs:90-113 */
void or_s(orc* v, int n, int q);                      /* single_qubit.rs:121-137 */
void or_t(orc* v, int n, int q);                      /* single_qubit.rs:146-161 */
void or_rx(orc* v, int n, int q, double theta);       /* parameterized.rs:48-78 */
void or_ry(orc* v, int n, int q, double theta);       /* parameterized.rs:101-131 */
void or_rz(orc* v, int n, int q, double theta);       /* parameterized.rs:154-176 */
void or_phase(orc* v, int n, int q, double theta);    /* parameterized.rs:199-218 */
void or_u3(orc* v, int n, int q, double theta, double phi, double lam); /* parameterized.rs:249-282 */
void or_cnot(orc* v, int n, int c, int t);            /* two_qubit.rs:12-31 */
void or_swap(orc* v, int n, int q1, int q2);          /* two_qubit.rs:40-62 */
void or_cz(orc* v, int n, int c, int t);              /* two_qubit.rs:71-90 */
void or_cphase(orc* v, int n, int c, int t, double angle); /* two_qubit.rs:100-136 */
void or_crx(orc* v, int n, int c, int t, double theta);    /* parameterized.rs:311-343 */
void or_cry(orc* v, int n, int c, int t, double theta);    /* parameterized.rs:368-399 */
void or_crz(orc* v, int n, int c, int t, double theta);    /* parameterized.rs:424-451 */
void or_cp(orc* v, int n, int c, int t, double theta);     /* parameterized.rs:476-499 */
void or_rxx(orc* v, int n, int q1, int q2, double theta);  /* parameterized.rs:528-578 */
void or_ryy(orc* v, int n, int q1, int q2, double theta);  /* parameterized.rs:604-653 */
void or_rzz(orc* v, int n, int q1, int q2, double theta);  /* parameterized.rs:679-709 */
void or_toffoli(orc* v, int n, int c1, int c2, int t);     /* three_qubit.rs:13-35 */

/* MatrixGate::apply (src/gates/mod.rs:77-93): out-of-place dense matvec
 * (row i = sum_j M[i][j] * v[j], in column order), then normalize. */
void or_matrix_gate(orc* v, int n, const orc* m, uint64_t dim, orc* scratch);
/* Full 2^n x 2^n matrix of a dense builder: Circuit::add_single_qubit_gate
 * (src/circuits/single_qubit.rs:12-63) via tensor_product
 * (src/circuits/utils.rs:11-49). */
void or_kron_single(const orc m[4], int n, int target, orc* full);
/* Full matrix of Circuit::add_two_qubit_gate (src/circuits/two_qubits.rs:11-82). */
void or_expand_two(const orc m[16], int n, int q1, int q2, orc* full);
/* The same two operators applied without materialising the matrix: 2x2 (4x4)
 * block on the target bit(s), then normalize.  Equal to the dense form up to
 * the signs of zeros (the skipped terms are exact zeros). */
void or_matrix1(orc* v, int n, int q, const orc m[4]);
void or_matrix2(orc* v, int n, int q1, int q2, const orc m[16]);

/* Circuit::execute (src/circuits/mod.rs:352-375) over gate descriptors. */
int or_apply_gate(orc* v, int n, const lq_gate* g);
int or_execute(orc* v, int n, const lq_gate* gates, int count);

/* Dense matrices of src/gates/matrix.rs:14-161 (2x2, row-major). */
void or_matrix_of(int which, double theta, orc m[4]); /* which: LQ_GATE_X/Y/Z/H/S/T/RX/RY/RZ */

/* QFT circuit builders (src/algorithms/qft.rs:10-46, 96-105).  Return the
 * gate count; gates may be NULL to query it. */
int or_qft_circuit(int n, lq_gate* gates);
int or_inverse_qft_circuit(int n, lq_gate* gates);

/* PauliTerm::expectation (src/optimization/observable.rs:95-147), the
 * reference's own loop (row-bit Y phase). */
double or_pauli_term_expectation(const orc* v, int n, const uint8_t* paulis, double coeff);
/* PauliObservable::expectation (observable.rs:228-231). */
double or_pauli_sum_expectation(const orc* v, int n, const uint8_t* paulis, const double* coeffs, int num_terms);

/* Ansatz::apply on a template: zero state, bind params, execute
 * (src/optimization/ansatz.rs:21-37). */
void or_run_template(orc* v, int n, const lq_gate* tmpl, int num_gates, const double* params, lq_gate* scratch);
/* ParameterShift::compute_gradient (src/optimization/gradient.rs:45-78). */
void or_param_shift(int n, const lq_gate* tmpl, int num_gates, const double* params, int num_params, double shift,
                    const uint8_t* paulis, const double* coeffs, int num_terms, double* grad_out);
/* The same loop over components first .. last-1 only (grad_out[i] for those
 * i); re-entrant, so disjoint ranges may run in parallel threads. */
void or_param_shift_range(int n, const lq_gate* tmpl, int num_gates, const double* params, int num_params,
                          double shift, const uint8_t* paulis, const double* coeffs, int num_terms, int first,
                          int last, double* grad_out);
/* FiniteDifference::compute_gradient (src/optimization/gradient.rs:99-125). */
void or_finite_difference(int n, const lq_gate* tmpl, int num_gates, const double* params, int num_params,
                          double eps, const uint8_t* paulis, const double* coeffs, int num_terms,
                          double* grad_out);
/* The energy E(params) = observable on the ansatz state (vqe.rs:63-67). */
double or_energy(int n, const lq_gate* tmpl, int num_gates, const double* params, const uint8_t* paulis,
                 const double* coeffs, int num_terms);

/* Optimizers (src/optimization/optimizer.rs:56-223) with callbacks. */
typedef double (*or_objective_fn)(const double* x, int dim, void* user);
typedef void (*or_gradient_fn)(const double* x, int dim, double* grad, void* user);
typedef struct {
    double optimal_value;
    int num_iterations; /* = length of the convergence history */
    int converged;
} or_opt_result;
/* GradientDescent::minimize (optimizer.rs:56-107).  x holds the initial
 * point on entry and the optimal parameters on exit; history (nullable,
 * capacity max_iterations) receives the convergence history. */
or_opt_result or_gd_minimize(double lr, int max_iterations, double tol, or_objective_fn f, or_gradient_fn g,
                             void* user, double* x, int dim, double* history);
/* Adam::minimize (optimizer.rs:139-223). */
or_opt_result or_adam_minimize(double lr, double beta1, double beta2, double eps, int max_iterations, double tol,
                               or_objective_fn f, or_gradient_fn g, void* user, double* x, int dim,
                               double* history);

/* VQE::run (src/optimization/vqe.rs:55-99) with ParameterShift and Adam
 * (optimizer 1) or GradientDescent (optimizer 0).  params: initial on entry,
 * optimal on exit.  final_state (nullable, 2^n amplitudes). */
or_opt_result or_vqe_run(int n, const lq_gate* tmpl, int num_gates, int num_params, const uint8_t* paulis,
                         const double* coeffs, int num_terms, double shift, int optimizer, double lr,
                         int max_iterations, double tol, double* params, double* history, orc* final_state);

/* calculate_energy_efficient (src/algorithms/xyz_heisenberg.rs:143-162):
 * open chain, +J sign. */
double or_xyz_energy_efficient(const orc* v, int n, double jx, double jy, double jz, double field);

/* Timing helper for the CPU baseline: monotonic seconds. */
double or_now(void);

#ifdef __cplusplus
}
#endif

#endif
