// Noise channels of Circuit::execute (src/noise/mod.rs:10-380,
// src/circuits/mod.rs:359-371) on the device state.
//
// The reference's noise loops address bit `1 << qubit` of the basis index
// (mod.rs:129, 197, 343-376): the low-order bit for qubit 0, where the gates
// use bit n-1-qubit.  The entry points here take the reference's qubit
// numbers and keep that bit order.
//
//  - amplitude damping (mod.rs:110-147): for every listed qubit q the sum of
//    both Kraus operators, (K0 + K1) psi, is accumulated into one output
//    from the same input psi, then State::normalize.  One pass per qubit:
//    a pair (i0 = bit q clear, i1 = i0 | 2^q) adds psi[i0] then
//    sqrt(gamma) psi[i1] to out[i0] and sqrt(1-gamma) psi[i1] to out[i1], in
//    the reference's order of additions.
//  - phase damping (mod.rs:173-206): the output is reset from psi for every
//    listed qubit, so only the last one acts: amplitudes with its bit set
//    are scaled by sqrt(1-lambda) (the diagonal kernel), then normalize.
//  - an empty qubit list leaves the output at zero for both (the vector is
//    zero and normalize keeps it).
//  - depolarizing (mod.rs:63-86): per listed qubit a uniform u < p draws an
//    error type, X, the reference's Y form (i psi[i^m] / -i psi[i^m], the
//    YGate's -Y) or Z, applied with the per-gate kernels.  The reference
//    draws from thread_rng; here uniforms come from a caller-held SplitMix64
//    state, so runs are reproducible.
#include <cmath>
#include <vector>

#include "lq_common.h"

namespace lq {

// A product rounded on its own (no contraction into the following add), as
// the reference's separate complex multiply and add round.
__device__ __forceinline__ double mul_rn(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ float mul_rn(float a, float b) { return __fmul_rn(a, b); }

template <typename V, typename T>
__global__ void __launch_bounds__(256) k_amp_damp_acc(const V* __restrict__ psi, V* __restrict__ out,
                                                      uint64_t pairs, int bit, T s1, T sg, int first) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t m = 1ull << bit;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < pairs; k += stride) {
        const uint64_t i0 = insert_zero(k, bit), i1 = i0 | m;
        const V a0 = psi[i0], a1 = psi[i1];
        V o0, o1;
        if (first) {
            o0.x = (T)0;
            o0.y = (T)0;
            o1 = o0;
        } else {
            o0 = out[i0];
            o1 = out[i1];
        }
        // new[i0] += psi[i0]; new[i0] += psi[i1] * sqrt(gamma)   (mod.rs:131-140)
        // new[i1] += psi[i1] * sqrt(1 - gamma)
        o0.x = o0.x + a0.x;
        o0.y = o0.y + a0.y;
        o0.x = o0.x + mul_rn(a1.x, sg);
        o0.y = o0.y + mul_rn(a1.y, sg);
        o1.x = o1.x + mul_rn(a1.x, s1);
        o1.y = o1.y + mul_rn(a1.y, s1);
        out[i0] = o0;
        out[i1] = o1;
    }
}

static unsigned grid_pairs(uint64_t pairs) {
    uint64_t b = (pairs + 255) / 256;
    if (b > 32768) b = 32768;
    return (unsigned)(b < 1 ? 1 : b);
}

static int check_noise_qubits(const lq_state* st, const int* qubits, int count) {
    if (count < 0 || (count > 0 && !qubits)) {
        set_error("invalid qubit list");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    for (int i = 0; i < count; ++i)
        if (qubits[i] < 0 || qubits[i] >= st->n) {
            set_error("Invalid qubit index: %d (must be < %d)", qubits[i], st->n);
            return LQ_ERR_INVALID_QUBIT_INDEX;
        }
    return LQ_OK;
}

static int check_probability(double p, const char* what) {
    if (!(p >= 0.0 && p <= 1.0)) {
        set_error("%s must lie in [0, 1] (got %g)", what, p);
        return LQ_ERR_INVALID_ARGUMENT;
    }
    return LQ_OK;
}

static int zero_state(lq_state* st) {
    LQ_HIP_CHECK(hipMemsetAsync(st->d_amps, 0, st->dim * st->amp_bytes, stream()));
    return LQ_OK;
}

template <typename V, typename T>
static int amp_damp_t(lq_state* st, const int* qubits, int count, double gamma) {
    V* out = (V*)scratch2(st->dim * sizeof(V));
    if (!out) {
        set_error("scratch allocation failed");
        return LQ_ERR_OUT_OF_MEMORY;
    }
    const T s1 = (T)std::sqrt(1.0 - gamma), sg = (T)std::sqrt(gamma);
    const uint64_t pairs = st->dim >> 1;
    for (int i = 0; i < count; ++i) {
        hipLaunchKernelGGL((k_amp_damp_acc<V, T>), dim3(grid_pairs(pairs)), dim3(256), 0, stream(),
                           (const V*)st->d_amps, out, pairs, qubits[i], s1, sg, i == 0 ? 1 : 0);
        LQ_LAUNCH_CHECK();
    }
    LQ_HIP_CHECK(hipMemcpyAsync(st->d_amps, out, st->dim * sizeof(V), hipMemcpyDeviceToDevice, stream()));
    return LQ_OK;
}

static uint64_t splitmix_next(uint64_t& s) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double uniform53(uint64_t& s) { return (double)(splitmix_next(s) >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace lq

using namespace lq;

extern "C" {

int lq_noise_amplitude_damping(lq_state* st, const int* qubits, int count, double gamma) {
    LQ_API_GUARD;
    clear_error();
    int rc = check_state(st);
    if (!rc) rc = check_noise_qubits(st, qubits, count);
    if (!rc) rc = check_probability(gamma, "gamma");
    if (rc) return rc;
    if (count == 0) {
        rc = zero_state(st);
    } else if (st->dtype == LQ_C128) {
        rc = amp_damp_t<double2, double>(st, qubits, count, gamma);
    } else {
        rc = amp_damp_t<float2, float>(st, qubits, count, gamma);
    }
    if (rc) return rc;
    return normalize_device(st, nullptr);
}

int lq_noise_phase_damping(lq_state* st, const int* qubits, int count, double lambda) {
    LQ_API_GUARD;
    clear_error();
    int rc = check_state(st);
    if (!rc) rc = check_noise_qubits(st, qubits, count);
    if (!rc) rc = check_probability(lambda, "lambda");
    if (rc) return rc;
    if (count == 0) {
        rc = zero_state(st);
    } else {
        const double d[8] = {std::sqrt(1.0 - lambda), 0.0, 1, 0, 1, 0, 1, 0};
        rc = launch_diag(st, 1ull << qubits[count - 1], -1, -1, d);
    }
    if (rc) return rc;
    return normalize_device(st, nullptr);
}

int lq_noise_depolarizing(lq_state* st, const int* qubits, int count, double p, uint64_t* rng_state,
                          int* errors) {
    LQ_API_GUARD;
    clear_error();
    int rc = check_state(st);
    if (!rc) rc = check_noise_qubits(st, qubits, count);
    if (!rc) rc = check_probability(p, "error probability");
    if (rc) return rc;
    if (!rng_state) {
        set_error("rng_state is null");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    uint64_t s = *rng_state;
    int applied = 0;
    for (int i = 0; i < count; ++i) {
        if (!(uniform53(s) < p)) continue;
        int e = (int)(uniform53(s) * 3.0);
        if (e > 2) e = 2;
        const int bit = qubits[i];
        static const double unused[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // X and Y need no matrix
        if (e == 0) {
            rc = launch_pair(st, bit, 0, unused, kOpSwap);
        } else if (e == 1) {
            rc = launch_pair(st, bit, 0, unused, kOpNegY);
        } else {
            const double d[8] = {-1.0, 0.0, 1, 0, 1, 0, 1, 0};
            rc = launch_diag(st, 1ull << bit, -1, -1, d);
        }
        if (rc) return rc;
        ++applied;
    }
    *rng_state = s;
    if (errors) *errors = applied;
    return LQ_OK;
}

}  // extern "C"
