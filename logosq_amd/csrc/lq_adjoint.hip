// Adjoint-method gradient of E(theta) = <psi|H|psi>, psi = U_G ... U_1 |0...0>.
// No counterpart in the reference (it offers ParameterShift and
// FiniteDifference, src/optimization/gradient.rs): one forward sweep, then
// lambda = H psi and a backward sweep.  For a rotation U_k = exp(-i theta/2 P)
// (RX, RY, RZ) dE/dtheta_k = Im <lambda_k| P |psi_k> with both states taken
// right after gate k; then U_k^dagger is applied to psi and lambda.  Equal to
// ParameterShift with shift pi/2 up to rounding, at the cost of ~3 state
// sweeps per gate instead of 2P circuit executions.  H uses the same Y-phase
// convention as the expectation (reference or physical).
#include <cmath>
#include <cstring>
#include <vector>

#include "lq_fused.h"

namespace lq {

void bind_template(const lq_gate* tmpl, int num_gates, const double* params, int num_params,
                   std::vector<lq_gate>& out);
int check_template(int n, const lq_gate* tmpl, int num_gates, int num_params);
int execute_fused(lq_state* st, const lq_gate* gates, int count);

constexpr int kAThreads = 256;

template <typename V>
__global__ void k_zero(V* __restrict__ a, uint64_t dim) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
        V z;
        z.x = 0;
        z.y = 0;
        a[i] = z;
    }
}

// lambda_i += w_x(i) psi_{i ^ x}, w_x(i) = sum_t (cr_t, ci_t) (-1)^{popc(i & zy_t)}
// (the matrix element <i|H|i^x> of the group's terms).
template <typename V>
__global__ void __launch_bounds__(kAThreads) k_pauli_apply_add(const V* __restrict__ psi, V* __restrict__ lam,
                                                               uint64_t dim, uint64_t x,
                                                               const PauliTermDev* __restrict__ terms, int nt) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
        double wr = 0.0, wi = 0.0;
        for (int t = 0; t < nt; ++t) {
            bool s = __popcll(i & terms[t].zy) & 1;
            wr += s ? -terms[t].cr : terms[t].cr;
            wi += s ? -terms[t].ci : terms[t].ci;
        }
        V b = psi[i ^ x];
        V l = lam[i];
        double br = (double)b.x, bi = (double)b.y;
        l.x = (decltype(l.x))((double)l.x + (wr * br - wi * bi));
        l.y = (decltype(l.y))((double)l.y + (wr * bi + wi * br));
        lam[i] = l;
    }
}

template <int NT>
__device__ __forceinline__ double bsum_a(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    return sh[0];
}

// Partial sums of Im <lambda| P_b |psi> for a single-qubit Pauli P on bit b:
// (X psi)_i = psi_{i^m}; (Y psi)_i = (bit_i ? i : -i) psi_{i^m}; (Z psi)_i = (bit_i ? -1 : 1) psi_i.
template <typename V>
__global__ void __launch_bounds__(kAThreads) k_pauli_im_inner(const V* __restrict__ lam, const V* __restrict__ psi,
                                                              uint64_t dim, int bit, int pauli,
                                                              double* __restrict__ partial) {
    __shared__ double sh[kAThreads];
    const uint64_t m = 1ull << bit;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    double acc = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
        const bool b = (i & m) != 0;
        V l = lam[i];
        V p = pauli == LQ_PAULI_Z ? psi[i] : psi[i ^ m];
        double pr = p.x, pi = p.y;
        if (pauli == LQ_PAULI_Z) {
            if (b) {
                pr = -pr;
                pi = -pi;
            }
        } else if (pauli == LQ_PAULI_Y) {
            // multiply by +i (bit set) or -i (bit clear)
            double r = pr;
            if (b) {
                pr = -pi;
                pi = r;
            } else {
                pr = pi;
                pi = -r;
            }
        }
        // Im(conj(l) * p) = l.x * p.y - l.y * p.x
        acc += (double)l.x * pi - (double)l.y * pr;
    }
    double t = bsum_a<kAThreads>(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ void __launch_bounds__(1024) k_sum_to(const double* __restrict__ partial, int count, double* __restrict__ out) {
    __shared__ double sh[1024];
    double s = 0.0;
    for (int i = threadIdx.x; i < count; i += 1024) s += partial[i];
    double t = bsum_a<1024>(s, sh);
    if (threadIdx.x == 0) out[0] = t;
}

// Dedicated device buffer: the expectation and normalisation helpers use the
// shared scratch area in between.
static void* g_adj = nullptr;
static size_t g_adj_cap = 0;

static char* adj_buffer(size_t bytes) {
    if (bytes <= g_adj_cap) return (char*)g_adj;
    if (hipStreamSynchronize(stream()) != hipSuccess) return nullptr;
    if (g_adj) (void)hipFree(g_adj);
    g_adj = nullptr;
    g_adj_cap = 0;
    if (hipMalloc(&g_adj, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    g_adj_cap = bytes;
    return (char*)g_adj;
}

static unsigned grid_of(uint64_t work) {
    uint64_t g = (work + kAThreads - 1) / kAThreads;
    if (g > 8192) g = 8192;
    if (g < 1) g = 1;
    return (unsigned)g;
}

static int red_blocks_a(uint64_t dim) {
    uint64_t b = (dim + kAThreads * 8 - 1) / (kAThreads * 8);
    if (b < 1) b = 1;
    if (b > 1024) b = 1024;
    return (int)b;
}

// Inverse of a gate as a gate descriptor; false if not supported.
static bool inverse_gate(const lq_gate& g, lq_gate& inv) {
    inv = g;
    switch (g.kind) {
        case LQ_GATE_X: case LQ_GATE_Y: case LQ_GATE_Z: case LQ_GATE_H: case LQ_GATE_CNOT: case LQ_GATE_SWAP:
        case LQ_GATE_CZ: case LQ_GATE_TOFFOLI:
            return true;  // involutions (the reference's -Y squares to I too)
        case LQ_GATE_S:
            inv.kind = LQ_GATE_PHASE;
            inv.theta[0] = -M_PI / 2.0;
            return true;
        case LQ_GATE_T:
            inv.kind = LQ_GATE_PHASE;
            inv.theta[0] = -M_PI / 4.0;
            return true;
        case LQ_GATE_RX: case LQ_GATE_RY: case LQ_GATE_RZ: case LQ_GATE_PHASE: case LQ_GATE_CPHASE:
        case LQ_GATE_CRX: case LQ_GATE_CRY: case LQ_GATE_CRZ: case LQ_GATE_CP: case LQ_GATE_RXX: case LQ_GATE_RYY:
        case LQ_GATE_RZZ:
            inv.theta[0] = -g.theta[0];
            return true;
        case LQ_GATE_U3:
            // U3(t, p, l)^-1 = U3(-t, -l, -p)
            inv.theta[0] = -g.theta[0];
            inv.theta[1] = -g.theta[2];
            inv.theta[2] = -g.theta[1];
            return true;
        default:
            return false;
    }
}

template <typename V>
static int adjoint_t(int n, int dtype, const lq_gate* tmpl, int num_gates, const double* params, int num_params,
                     const PauliSum& ps, double* energy_out, double* grad_out) {
    std::vector<lq_gate> g;
    bind_template(tmpl, num_gates, params, num_params, g);
    for (int k = 0; k < num_gates; ++k) {
        lq_gate inv;
        if ((g[k].flags & LQ_GATE_FLAG_RENORMALIZE) || !inverse_gate(g[k], inv)) {
            set_error("adjoint gradient: gate %d (kind %d) has no unitary inverse here", k, g[k].kind);
            return LQ_ERR_UNSUPPORTED;
        }
        for (int a = 0; a < 3; ++a)
            if (tmpl[k].param[a] >= 0 && !(a == 0 && (g[k].kind == LQ_GATE_RX || g[k].kind == LQ_GATE_RY ||
                                                      g[k].kind == LQ_GATE_RZ))) {
                set_error("adjoint gradient supports parameters on RX / RY / RZ only (gate %d)", k);
                return LQ_ERR_UNSUPPORTED;
            }
    }
    lq_state *psi = nullptr, *lam = nullptr;
    int rc = lq_state_create(n, dtype, &psi);
    if (rc) return rc;
    rc = lq_state_create(n, dtype, &lam);
    if (rc) {
        lq_state_destroy(psi);
        return rc;
    }
    const uint64_t dim = 1ull << n;
    // device: terms, per-gate results, partials
    const int nb = red_blocks_a(dim);
    size_t tb = std::max<size_t>(ps.terms.size(), 1) * sizeof(PauliTermDev);
    size_t need = tb + (size_t)(num_gates + 2) * sizeof(double) + (size_t)nb * sizeof(double) + 1024;
    char* base = adj_buffer(need);
    if (!base) {
        lq_state_destroy(psi);
        lq_state_destroy(lam);
        set_error("scratch allocation failed");
        return LQ_ERR_OUT_OF_MEMORY;
    }
    PauliTermDev* d_terms = (PauliTermDev*)base;
    double* d_res = (double*)(base + ((tb + 255) & ~(size_t)255));
    double* d_part = d_res + num_gates + 2;
    std::vector<PauliTermDev> hterms = ps.terms;
    if (!hterms.empty()) {
        // pageable source: complete the copy before the vector can go away
        if (hipMemcpyAsync(d_terms, hterms.data(), hterms.size() * sizeof(PauliTermDev), hipMemcpyHostToDevice,
                           stream()) != hipSuccess ||
            hipStreamSynchronize(stream()) != hipSuccess) {
            set_error("term upload failed");
            rc = LQ_ERR_DEVICE;
        }
    }
    // forward
    if (!rc) rc = execute_fused(psi, g.data(), num_gates);
    if (!rc && energy_out) rc = expect_pauli_sum(psi, ps, energy_out);
    // lambda = H psi
    if (!rc) {
        hipLaunchKernelGGL(k_zero<V>, dim3(grid_of(dim)), dim3(kAThreads), 0, stream(), (V*)lam->d_amps, dim);
        for (const PauliGroup& grp : ps.groups)
            hipLaunchKernelGGL(k_pauli_apply_add<V>, dim3(grid_of(dim)), dim3(kAThreads), 0, stream(),
                               (const V*)psi->d_amps, (V*)lam->d_amps, dim, grp.x,
                               (const PauliTermDev*)(d_terms + grp.first), grp.count);
        if (hipGetLastError() != hipSuccess) {
            set_error("Pauli apply launch failed");
            rc = LQ_ERR_DEVICE;
        }
    }
    // backward
    std::vector<int> gate_param(num_gates, -1);
    for (int k = num_gates - 1; k >= 0 && !rc; --k) {
        if (tmpl[k].param[0] >= 0) {
            gate_param[k] = tmpl[k].param[0];
            int pauli = g[k].kind == LQ_GATE_RX ? LQ_PAULI_X : g[k].kind == LQ_GATE_RY ? LQ_PAULI_Y : LQ_PAULI_Z;
            int bit = n - 1 - g[k].qubits[0];
            hipLaunchKernelGGL(k_pauli_im_inner<V>, dim3(nb), dim3(kAThreads), 0, stream(), (const V*)lam->d_amps,
                               (const V*)psi->d_amps, dim, bit, pauli, d_part);
            hipLaunchKernelGGL(k_sum_to, dim3(1), dim3(1024), 0, stream(), (const double*)d_part, nb, d_res + k);
            if (hipGetLastError() != hipSuccess) {
                set_error("inner-product launch failed");
                rc = LQ_ERR_DEVICE;
                break;
            }
        }
        if (k == 0) break;  // the state before the first gate is not needed
        lq_gate inv;
        inverse_gate(g[k], inv);
        rc = apply_gate_desc(psi, &inv);
        if (!rc) rc = apply_gate_desc(lam, &inv);
    }
    std::vector<double> res(num_gates, 0.0);
    if (!rc && num_gates > 0) {
        if (hipMemcpyAsync(res.data(), d_res, num_gates * sizeof(double), hipMemcpyDeviceToHost, stream()) !=
                hipSuccess ||
            hipStreamSynchronize(stream()) != hipSuccess) {
            set_error("result copy failed");
            rc = LQ_ERR_DEVICE;
        }
    }
    if (!rc) {
        for (int i = 0; i < num_params; ++i) grad_out[i] = 0.0;
        for (int k = 0; k < num_gates; ++k)
            if (gate_param[k] >= 0) grad_out[gate_param[k]] += res[k];
    }
    lq_state_destroy(psi);
    lq_state_destroy(lam);
    return rc;
}

}  // namespace lq

using namespace lq;

extern "C" int lq_adjoint_gradient(int num_qubits, int dtype, const lq_gate* tmpl, int num_gates,
                                   const double* params, int num_params, const uint8_t* paulis, const double* coeffs,
                                   int num_terms, int convention, double* energy_out, double* grad_out) {
    clear_error();
    if (num_qubits < 1 || num_qubits > kMaxQubits || num_params < 0 || (num_params > 0 && (!params || !grad_out)) ||
        (dtype != LQ_C128 && dtype != LQ_C64)) {
        set_error("invalid arguments");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    int rc = check_template(num_qubits, tmpl, num_gates, num_params);
    if (rc) return rc;
    PauliSum ps;
    rc = compile_pauli_sum(num_qubits, paulis, coeffs, num_terms, convention, &ps);
    if (rc) return rc;
    if (dtype == LQ_C128)
        return adjoint_t<double2>(num_qubits, dtype, tmpl, num_gates, params, num_params, ps, energy_out, grad_out);
    return adjoint_t<float2>(num_qubits, dtype, tmpl, num_gates, params, num_params, ps, energy_out, grad_out);
}
