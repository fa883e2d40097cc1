// Classical optimisers and the VQE driver (host code): GradientDescent and
// Adam (src/optimization/optimizer.rs:35-224) and VQE::run
// (src/optimization/vqe.rs:55-99), with the objective and gradient evaluated
// by the GPU path.
#include <algorithm>
#include <cmath>
#include <vector>

#include "lq_host.h"

namespace lq {

// f64::powi for a run-time exponent: binary exponentiation (LLVM powi).
static double powi_f64(double a, int b) {
    bool recip = b < 0;
    unsigned e = recip ? (unsigned)(-b) : (unsigned)b;
    double r = 1.0;
    for (;;) {
        if (e & 1u) r *= a;
        e >>= 1;
        if (e == 0) break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}

static double l2(const std::vector<double>& g) {
    double s = 0.0;
    for (double v : g) s += v * v;
    return std::sqrt(s);
}

struct Cb {
    lq_objective_fn f;
    lq_gradient_fn g;
    void* user;
};

int gd_minimize(double lr, int max_it, double tol, const Cb& cb, double* x, int dim, double* history,
                lq_opt_result* res) {
    std::vector<double> grad(dim > 0 ? dim : 1);
    int hist = 0;
    bool converged = false;
    for (int it = 0; it < max_it; ++it) {
        double value = cb.f(x, dim, cb.user);
        if (history) history[hist] = value;
        ++hist;
        cb.g(x, dim, grad.data(), cb.user);
        if (l2(grad) < tol) {
            converged = true;
            break;
        }
        for (int i = 0; i < dim; ++i) x[i] -= lr * grad[i];
    }
    res->optimal_value = cb.f(x, dim, cb.user);
    cb.g(x, dim, grad.data(), cb.user);
    if (l2(grad) < tol) converged = true;
    res->num_iterations = hist;
    res->converged = converged ? 1 : 0;
    return LQ_OK;
}

// target: stop as soon as a recorded value is at or below it (-inf: never,
// Adam::minimize itself); *reached (nullable) tells whether it was.
int adam_minimize(double lr, double b1, double b2, double eps, int max_it, double tol, const Cb& cb, double* x,
                  int dim, double* history, lq_opt_result* res, double target = -INFINITY, bool* reached = nullptr) {
    std::vector<double> m(dim, 0.0), v(dim, 0.0), grad(dim > 0 ? dim : 1);
    int hist = 0;
    bool converged = false, have_prev = false;
    double prev = 0.0;
    if (reached) *reached = false;
    for (int it = 1; it <= max_it; ++it) {
        double value = cb.f(x, dim, cb.user);
        if (history) history[hist] = value;
        ++hist;
        if (value <= target) {
            if (reached) *reached = true;
            break;
        }
        if (have_prev && std::fabs(prev - value) < tol) {
            converged = true;
            break;
        }
        prev = value;
        have_prev = true;
        cb.g(x, dim, grad.data(), cb.user);
        if (l2(grad) < tol) {
            converged = true;
            break;
        }
        for (int i = 0; i < dim; ++i) {
            m[i] = b1 * m[i] + (1.0 - b1) * grad[i];
            v[i] = b2 * v[i] + (1.0 - b2) * grad[i] * grad[i];
        }
        double mc = 1.0 / (1.0 - powi_f64(b1, it));
        double vc = 1.0 / (1.0 - powi_f64(b2, it));
        double sn = 0.0;
        for (int i = 0; i < dim; ++i) {
            double mh = m[i] * mc, vh = v[i] * vc;
            double step = lr * mh / (std::sqrt(vh) + eps);
            x[i] -= step;
            sn += step * step;
        }
        if (std::sqrt(sn) < tol) {
            converged = true;
            break;
        }
    }
    res->optimal_value = cb.f(x, dim, cb.user);
    cb.g(x, dim, grad.data(), cb.user);
    if (l2(grad) < tol) converged = true;
    res->num_iterations = hist;
    res->converged = converged ? 1 : 0;
    return LQ_OK;
}

// VQE objective / gradient on the GPU path.
struct VqeCtx {
    int n, dtype, num_gates, num_params, grad_mode;
    const lq_gate* tmpl;
    const PauliSum* ps;
    double shift;
    lq_state* st;
    std::vector<lq_gate> sg;
    const uint8_t* paulis;
    const double* coeffs;
    int num_terms, convention;
    int rc;
    // adjoint / batched modes: the gradient evaluated together with the
    // energy in vqe_f, kept for the vqe_g call at the same point
    std::vector<double> cached_x, cached_g;
    bool cached = false;
};

int energy_sequential(lq_state* st, const lq_gate* tmpl, int num_gates, const double* params, int num_params,
                      const PauliSum& ps, std::vector<lq_gate>& scratch_gates, double* out);

// Sequential mode evaluates E with per-gate kernels, as the reference's
// VQE does (vqe.rs:63-70).  The batched and adjoint gradients produce
// E(x) together with the gradient, so there one call serves both the
// objective and the gradient request that Adam / GD make at the same x.
static double vqe_f(const double* x, int dim, void* user) {
    VqeCtx* c = (VqeCtx*)user;
    double e = 0.0;
    if (c->rc) return NAN;
    if (c->grad_mode == LQ_GRAD_ADJOINT || c->grad_mode == LQ_GRAD_BATCHED) {
        c->cached_g.assign(dim > 0 ? dim : 1, 0.0);
        if (c->grad_mode == LQ_GRAD_ADJOINT)
            c->rc = lq_adjoint_gradient(c->n, c->dtype, c->tmpl, c->num_gates, x, dim, c->paulis, c->coeffs,
                                        c->num_terms, c->convention, &e, c->cached_g.data());
        else
            c->rc = lq_param_shift_gradient(c->n, c->dtype, c->tmpl, c->num_gates, x, dim, c->shift, c->paulis,
                                            c->coeffs, c->num_terms, c->convention, LQ_GRAD_BATCHED, &e,
                                            c->cached_g.data());
        c->cached_x.assign(x, x + dim);
        c->cached = c->rc == 0;
        return c->rc ? NAN : e;
    }
    c->rc = energy_sequential(c->st, c->tmpl, c->num_gates, x, dim, *c->ps, c->sg, &e);
    return c->rc ? NAN : e;
}

static void vqe_g(const double* x, int dim, double* grad, void* user) {
    VqeCtx* c = (VqeCtx*)user;
    if (c->rc) {
        for (int i = 0; i < dim; ++i) grad[i] = NAN;
        return;
    }
    if (c->cached && (int)c->cached_x.size() == dim && std::equal(x, x + dim, c->cached_x.begin())) {
        for (int i = 0; i < dim; ++i) grad[i] = c->cached_g[i];
        return;
    }
    if (c->grad_mode == LQ_GRAD_ADJOINT)
        c->rc = lq_adjoint_gradient(c->n, c->dtype, c->tmpl, c->num_gates, x, dim, c->paulis, c->coeffs,
                                    c->num_terms, c->convention, nullptr, grad);
    else
        c->rc = lq_param_shift_gradient(c->n, c->dtype, c->tmpl, c->num_gates, x, dim, c->shift, c->paulis,
                                        c->coeffs, c->num_terms, c->convention, c->grad_mode, nullptr, grad);
    if (c->rc)
        for (int i = 0; i < dim; ++i) grad[i] = NAN;
}

int check_template(int n, const lq_gate* tmpl, int num_gates, int num_params);

}  // namespace lq

using namespace lq;

extern "C" {

int lq_gd_minimize(double learning_rate, int max_iterations, double tolerance, lq_objective_fn f, lq_gradient_fn g,
                   void* user, double* x, int dim, double* history, lq_opt_result* result) {
    LQ_API_GUARD;
    clear_error();
    if (!f || !g || !result || dim < 0 || (dim > 0 && !x)) {
        set_error("invalid optimiser arguments");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    Cb cb{f, g, user};
    return gd_minimize(learning_rate, max_iterations, tolerance, cb, x, dim, history, result);
}

int lq_adam_minimize(double learning_rate, double beta1, double beta2, double epsilon, int max_iterations,
                     double tolerance, lq_objective_fn f, lq_gradient_fn g, void* user, double* x, int dim,
                     double* history, lq_opt_result* result) {
    LQ_API_GUARD;
    clear_error();
    if (!f || !g || !result || dim < 0 || (dim > 0 && !x)) {
        set_error("invalid optimiser arguments");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    Cb cb{f, g, user};
    return adam_minimize(learning_rate, beta1, beta2, epsilon, max_iterations, tolerance, cb, x, dim, history,
                         result);
}

static int vqe_run(int num_qubits, int dtype, const lq_gate* tmpl, int num_gates, int num_params,
                   const uint8_t* paulis, const double* coeffs, int num_terms, int convention, double shift,
                   int grad_mode, int optimizer, double learning_rate, int max_iterations, double tolerance,
                   double* params, double* history, lq_state* final_state, lq_opt_result* result, double target,
                   bool* reached) {
    clear_error();
    if (!result || num_params < 0 || (num_params > 0 && !params)) {
        set_error("invalid VQE arguments");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    int rc = check_template(num_qubits, tmpl, num_gates, num_params);
    if (rc) return rc;
    PauliSum ps;
    rc = compile_pauli_sum(num_qubits, paulis, coeffs, num_terms, convention, &ps);
    if (rc) return rc;
    VqeCtx c;
    c.n = num_qubits;
    c.dtype = dtype;
    c.num_gates = num_gates;
    c.num_params = num_params;
    c.grad_mode = grad_mode;
    c.tmpl = tmpl;
    c.ps = &ps;
    c.shift = shift;
    c.paulis = paulis;
    c.coeffs = coeffs;
    c.num_terms = num_terms;
    c.convention = convention;
    c.rc = 0;
    c.st = nullptr;
    rc = lq_state_create(num_qubits, dtype, &c.st);
    if (rc) return rc;
    Cb cb{vqe_f, vqe_g, &c};
    if (optimizer == 1)
        rc = adam_minimize(learning_rate, 0.9, 0.999, 1e-8, max_iterations, tolerance, cb, params, num_params, history,
                           result, target, reached);
    else
        rc = gd_minimize(learning_rate, max_iterations, tolerance, cb, params, num_params, history, result);
    if (!rc) rc = c.rc;
    if (!rc && final_state) {
        if (final_state->n != num_qubits || final_state->dtype != dtype) {
            set_error("final_state has the wrong shape");
            rc = LQ_ERR_DIMENSION_MISMATCH;
        } else {
            double e = 0.0;
            rc = energy_sequential(final_state, tmpl, num_gates, params, num_params, ps, c.sg, &e);
        }
    }
    lq_state_destroy(c.st);
    return rc;
}

int lq_vqe_run(int num_qubits, int dtype, const lq_gate* tmpl, int num_gates, int num_params,
               const uint8_t* paulis, const double* coeffs, int num_terms, int convention, double shift,
               int grad_mode, int optimizer, double learning_rate, int max_iterations, double tolerance,
               double* params, double* history, lq_state* final_state, lq_opt_result* result) {
    LQ_API_GUARD;
    return vqe_run(num_qubits, dtype, tmpl, num_gates, num_params, paulis, coeffs, num_terms, convention, shift,
                   grad_mode, optimizer, learning_rate, max_iterations, tolerance, params, history, final_state,
                   result, -INFINITY, nullptr);
}

int lq_vqe_run_to_target(int num_qubits, int dtype, const lq_gate* tmpl, int num_gates, int num_params,
                         const uint8_t* paulis, const double* coeffs, int num_terms, int convention, double shift,
                         int grad_mode, double learning_rate, int max_iterations, double tolerance,
                         double target_energy, double* params, double* history, lq_opt_result* result,
                         int32_t* reached) {
    LQ_API_GUARD;
    bool r = false;
    int rc = vqe_run(num_qubits, dtype, tmpl, num_gates, num_params, paulis, coeffs, num_terms, convention, shift,
                     grad_mode, 1, learning_rate, max_iterations, tolerance, params, history, nullptr, result,
                     target_energy, &r);
    if (reached) *reached = r ? 1 : 0;
    return rc;
}

int lq_vqe_adam_batch(int num_qubits, int dtype, const lq_gate* tmpl, int num_gates, int num_params,
                      const uint8_t* paulis, const double* coeffs, int num_terms, int convention, int count,
                      double learning_rate, int max_iterations, double tolerance, double* params, double* history,
                      double* energies, int32_t* iterations, int32_t* converged) {
    LQ_API_GUARD;
    clear_error();
    if (count < 0 || num_params < 0 || max_iterations < 0 || (count > 0 && (!params || !energies || !iterations ||
                                                                            !converged || !coeffs))) {
        set_error("invalid VQE batch arguments");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    int rc = check_template(num_qubits, tmpl, num_gates, num_params);
    if (rc) return rc;
    const double b1 = 0.9, b2 = 0.999, eps = 1e-8;
    const size_t P = (size_t)num_params;
    // per instance: Adam moments, history length, previous value, iteration,
    // and phase 0 = iterating, 1 = final evaluation pending, 2 = done
    std::vector<double> m((size_t)count * P, 0.0), v((size_t)count * P, 0.0), prev(count, 0.0);
    std::vector<int> hist(count, 0), it(count, 1), phase(count, max_iterations > 0 ? 0 : 1);
    std::vector<char> have_prev(count, 0), conv(count, 0);
    std::vector<int> act;
    std::vector<double> xb, cb, eb, gb;
    for (;;) {
        act.clear();
        for (int i = 0; i < count; ++i)
            if (phase[i] < 2) act.push_back(i);
        if (act.empty()) break;
        const int na = (int)act.size();
        xb.resize((size_t)na * P + 1);
        cb.resize((size_t)na * num_terms + 1);
        eb.resize(na);
        gb.resize((size_t)na * P + 1);
        for (int a = 0; a < na; ++a) {
            std::copy(params + (size_t)act[a] * P, params + (size_t)act[a] * P + P, xb.begin() + (size_t)a * P);
            std::copy(coeffs + (size_t)act[a] * num_terms, coeffs + (size_t)act[a] * num_terms + num_terms,
                      cb.begin() + (size_t)a * num_terms);
        }
        rc = lq_adjoint_gradient_batch(num_qubits, dtype, tmpl, num_gates, na, xb.data(), num_params, paulis,
                                       cb.data(), num_terms, convention, eb.data(), gb.data());
        if (rc) return rc;
        for (int a = 0; a < na; ++a) {
            const int i = act[a];
            double* x = params + (size_t)i * P;
            const double* g = gb.data() + (size_t)a * P;
            const double value = eb[a];
            double gn = 0.0;
            for (size_t k = 0; k < P; ++k) gn += g[k] * g[k];
            gn = std::sqrt(gn);
            if (phase[i] == 1) {  // the evaluation after the loop (optimizer.rs:208-214)
                energies[i] = value;
                if (gn < tolerance) conv[i] = 1;
                phase[i] = 2;
                continue;
            }
            if (history) history[(size_t)i * max_iterations + hist[i]] = value;
            ++hist[i];
            // the loop left at this x: the final evaluation is this one
            bool stop = (have_prev[i] && std::fabs(prev[i] - value) < tolerance);
            if (!stop) {
                prev[i] = value;
                have_prev[i] = 1;
                stop = gn < tolerance;
            }
            if (stop) {
                conv[i] = 1;
                energies[i] = value;
                phase[i] = 2;
                continue;
            }
            double* mi = m.data() + (size_t)i * P;
            double* vi = v.data() + (size_t)i * P;
            for (size_t k = 0; k < P; ++k) {
                mi[k] = b1 * mi[k] + (1.0 - b1) * g[k];
                vi[k] = b2 * vi[k] + (1.0 - b2) * g[k] * g[k];
            }
            const double mc = 1.0 / (1.0 - powi_f64(b1, it[i]));
            const double vc = 1.0 / (1.0 - powi_f64(b2, it[i]));
            double sn = 0.0;
            for (size_t k = 0; k < P; ++k) {
                const double step = learning_rate * (mi[k] * mc) / (std::sqrt(vi[k] * vc) + eps);
                x[k] -= step;
                sn += step * step;
            }
            if (std::sqrt(sn) < tolerance) {
                conv[i] = 1;
                phase[i] = 1;
            } else if (it[i] == max_iterations) {
                phase[i] = 1;
            } else {
                ++it[i];
            }
        }
    }
    for (int i = 0; i < count; ++i) {
        iterations[i] = hist[i];
        converged[i] = conv[i];
    }
    return LQ_OK;
}

}  // extern "C"
