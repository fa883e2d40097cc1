// Parameter-shift gradient over an ansatz template: the replacement of
// ParameterShift::compute_gradient (src/optimization/gradient.rs:45-78) and of
// the VQE objective closure (src/optimization/vqe.rs:63-67).
#include <cstring>
#include <vector>

#include "lq_host.h"

namespace lq {

// Bind parameter slots: theta[k] = params[param[k]] where param[k] >= 0.
void bind_template(const lq_gate* tmpl, int num_gates, const double* params, int num_params,
                   std::vector<lq_gate>& out) {
    out.assign(tmpl, tmpl + num_gates);
    for (auto& g : out)
        for (int k = 0; k < 3; ++k)
            if (g.param[k] >= 0 && g.param[k] < num_params) g.theta[k] = params[g.param[k]];
}

int check_template(int n, const lq_gate* tmpl, int num_gates, int num_params) {
    if (num_gates < 0 || (num_gates > 0 && !tmpl)) {
        set_error("invalid template");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    for (int i = 0; i < num_gates; ++i) {
        int rc = validate_gate(n, &tmpl[i]);
        if (rc) return rc;
        for (int k = 0; k < 3; ++k)
            if (tmpl[i].param[k] >= num_params) {
                set_error("Parameter count mismatch: gate %d uses slot %d of %d", i, tmpl[i].param[k], num_params);
                return LQ_ERR_PARAMETER_COUNT_MISMATCH;
            }
    }
    return LQ_OK;
}

// E(params): |0..0>, execute (per-gate kernels, reference order), expectation.
int energy_sequential(lq_state* st, const lq_gate* tmpl, int num_gates, const double* params, int num_params,
                      const PauliSum& ps, std::vector<lq_gate>& scratch_gates, double* out) {
    int rc = lq_state_set_basis(st, 0);
    if (rc) return rc;
    bind_template(tmpl, num_gates, params, num_params, scratch_gates);
    for (int g = 0; g < num_gates; ++g) {
        rc = apply_gate_desc(st, &scratch_gates[g]);
        if (rc) return rc;
    }
    return expect_pauli_sum(st, ps, out);
}

int param_shift_batched(int n, int dtype, const lq_gate* tmpl, int num_gates, const double* params, int num_params,
                        double shift, const PauliSum& ps, double* energy_out, double* grad_out);

}  // namespace lq

using namespace lq;

extern "C" int lq_param_shift_gradient(int num_qubits, int dtype, const lq_gate* tmpl, int num_gates,
                                       const double* params, int num_params, double shift, const uint8_t* paulis,
                                       const double* coeffs, int num_terms, int convention, int mode,
                                       double* energy_out, double* grad_out) {
    LQ_API_GUARD;
    clear_error();
    if (num_qubits < 1 || num_qubits > kMaxQubits || num_params < 0 || (num_params > 0 && (!params || !grad_out))) {
        set_error("invalid arguments");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    int rc = check_template(num_qubits, tmpl, num_gates, num_params);
    if (rc) return rc;
    PauliSum ps;
    rc = compile_pauli_sum(num_qubits, paulis, coeffs, num_terms, convention, &ps);
    if (rc) return rc;
    if (mode == LQ_GRAD_BATCHED)
        return param_shift_batched(num_qubits, dtype, tmpl, num_gates, params, num_params, shift, ps, energy_out,
                                   grad_out);
    if (mode != LQ_GRAD_SEQUENTIAL) {
        set_error("unknown gradient mode %d", mode);
        return LQ_ERR_INVALID_ARGUMENT;
    }
    lq_state* st = nullptr;
    rc = lq_state_create(num_qubits, dtype, &st);
    if (rc) return rc;
    std::vector<lq_gate> sg;
    std::vector<double> p(params, params + num_params);
    if (energy_out) rc = energy_sequential(st, tmpl, num_gates, params, num_params, ps, sg, energy_out);
    // gradient.rs:53-74: for each i, shift +s, run, expectation; -s, run, expectation
    for (int i = 0; i < num_params && !rc; ++i) {
        double ep = 0.0, em = 0.0;
        p[i] = params[i] + shift;
        rc = energy_sequential(st, tmpl, num_gates, p.data(), num_params, ps, sg, &ep);
        if (rc) break;
        p[i] = params[i] - shift;
        rc = energy_sequential(st, tmpl, num_gates, p.data(), num_params, ps, sg, &em);
        p[i] = params[i];
        grad_out[i] = (ep - em) / 2.0;
    }
    lq_state_destroy(st);
    return rc;
}
