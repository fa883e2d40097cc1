// C ABI for gates and circuits: validates descriptors the way the reference's
// builders do and dispatches each gate kind to its kernel.
#include <cmath>
#include <cstring>
#include <vector>

#include "lq_common.h"
#include "lq_fused.h"

namespace lq {

static inline uint64_t bitmask(int n, int q) { return 1ull << (n - 1 - q); }
static inline int bitpos(int n, int q) { return n - 1 - q; }

static void polar(double r, double th, double* out) {
    // Complex64::from_polar (r cos th, r sin th)
    out[0] = r * std::cos(th);
    out[1] = r * std::sin(th);
}

int gate_arity(int kind) {
    switch (kind) {
        case LQ_GATE_X: case LQ_GATE_Y: case LQ_GATE_Z: case LQ_GATE_H: case LQ_GATE_S: case LQ_GATE_T:
        case LQ_GATE_RX: case LQ_GATE_RY: case LQ_GATE_RZ: case LQ_GATE_PHASE: case LQ_GATE_U3:
        case LQ_GATE_MATRIX1:
            return 1;
        case LQ_GATE_CNOT: case LQ_GATE_SWAP: case LQ_GATE_CZ: case LQ_GATE_CPHASE: case LQ_GATE_CRX:
        case LQ_GATE_CRY: case LQ_GATE_CRZ: case LQ_GATE_CP: case LQ_GATE_RXX: case LQ_GATE_RYY: case LQ_GATE_RZZ:
        case LQ_GATE_MATRIX2:
            return 2;
        case LQ_GATE_TOFFOLI:
            return 3;
        default:
            return -1;
    }
}

int validate_gate(int n, const lq_gate* g) {
    int ar = gate_arity(g->kind);
    if (ar < 0) {
        set_error("unknown gate kind %d", g->kind);
        return LQ_ERR_INVALID_ARGUMENT;
    }
    for (int i = 0; i < ar; ++i) {
        int q = g->qubits[i];
        if (q < 0 || q >= n) {
            set_error("Invalid qubit index: %d (must be < %d)", q, n);
            return LQ_ERR_INVALID_QUBIT_INDEX;
        }
        for (int j = 0; j < i; ++j)
            if (g->qubits[j] == q) {
                // Circuit builders: "Control and target must be different" /
                // "Qubits must be different" (src/circuits/two_qubits.rs:54,217)
                set_error("Qubits must be different (gate kind %d uses qubit %d twice)", g->kind, q);
                return LQ_ERR_INVALID_ARGUMENT;
            }
    }
    return LQ_OK;
}

// 2x2 matrices of the rotation gates (host, f64, the reference's formulas):
// RX/CRX (parameterized.rs:70-74), RY/CRY (123-127), U3 (254-261).
void gate_pair_matrix(int kind, const double* th, double* m) {
    if (kind == LQ_GATE_RX || kind == LQ_GATE_CRX) {
        double c = std::cos(th[0] / 2.0), s = std::sin(th[0] / 2.0);
        double r[8] = {c, 0, 0, -s, 0, -s, c, 0};
        std::memcpy(m, r, sizeof(r));
    } else if (kind == LQ_GATE_RY || kind == LQ_GATE_CRY) {
        double c = std::cos(th[0] / 2.0), s = std::sin(th[0] / 2.0);
        double r[8] = {c, 0, -s, 0, s, 0, c, 0};
        std::memcpy(m, r, sizeof(r));
    } else {  // U3
        double c = std::cos(th[0] / 2.0), s = std::sin(th[0] / 2.0);
        double p[2];
        m[0] = c;
        m[1] = 0.0;
        polar(s, th[2], p);
        m[2] = -p[0];
        m[3] = -p[1];
        polar(s, th[1], p);
        m[4] = p[0];
        m[5] = p[1];
        polar(c, th[1] + th[2], p);
        m[6] = p[0];
        m[7] = p[1];
    }
}

// Run primitives with the per-gate kernels, in order.
int run_prims_per_gate(lq_state* st, const std::vector<Prim>& prims) {
    for (const Prim& p : prims) {
        int rc;
        switch (p.kind) {
            case kTPairMat: rc = launch_pair(st, p.b0, p.ctrl, p.m, kOpMatrix); break;
            case kTPairSwap: rc = launch_pair(st, p.b0, p.ctrl, p.m, kOpSwap); break;
            case kTPairNegY: rc = launch_pair(st, p.b0, p.ctrl, p.m, kOpNegY); break;
            case kTPairH: rc = launch_pair(st, p.b0, p.ctrl, p.m, 3 /* Hadamard op */); break;
            case kTDiag: rc = launch_diag(st, p.ctrl, p.sel0, p.sel1, p.m); break;
            case kTSwap2: rc = launch_swap2(st, p.b0, p.b1); break;
            case kTQuad: rc = launch_quad(st, p.b0, p.b1, p.m); break;
            default: rc = normalize_device(st, nullptr); break;
        }
        if (rc) return rc;
    }
    return LQ_OK;
}

// Per-gate dispatch (Gate::apply).  Angles are taken from g->theta.
int apply_gate_desc(lq_state* st, const lq_gate* g) {
    std::vector<Prim> prims;
    int rc = gates_to_prims(st->n, g, 1, prims);
    if (rc) return rc;
    return run_prims_per_gate(st, prims);
}

// RXX / RYY as a 4x4 block with the reference's orientation rule: the
// 00<->11 pair always, the 10<->01 pair only when qubit1 > qubit2
// (parameterized.rs:546-577, 621-652).  Sub-index (bit(q1) << 1) | bit(q2).
void two_qubit_rot_matrix(int kind, double theta, int q1, int q2, double* M) {
    double c = std::cos(theta / 2.0), s = std::sin(theta / 2.0);
    std::memset(M, 0, 32 * sizeof(double));
    auto set = [&](int r, int col, double re, double im) {
        M[8 * r + 2 * col] = re;
        M[8 * r + 2 * col + 1] = im;
    };
    double o0011 = kind == LQ_GATE_RXX ? -s : s;  // imaginary part of the 00<->11 coupling
    set(0, 0, c, 0);
    set(0, 3, 0, o0011);
    set(3, 0, 0, o0011);
    set(3, 3, c, 0);
    if (q1 > q2) {
        set(2, 2, c, 0);
        set(2, 1, 0, -s);
        set(1, 2, 0, -s);
        set(1, 1, c, 0);
    } else {
        set(1, 1, 1, 0);
        set(2, 2, 1, 0);
    }
}

}  // namespace lq

using namespace lq;

extern "C" {

int lq_apply_gate(lq_state* st, const lq_gate* gate) {
    LQ_API_GUARD;
    clear_error();
    int rc = check_state(st);
    if (rc) return rc;
    if (!gate) {
        set_error("gate is null");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    return apply_gate_desc(st, gate);
}

int lq_apply_1q(lq_state* st, int qubit, const double m[8]) {
    LQ_API_GUARD;
    clear_error();
    int rc = check_state(st);
    if (rc) return rc;
    rc = check_qubit(st, qubit);
    if (rc) return rc;
    if (!m) {
        set_error("matrix is null");
        return LQ_ERR_INVALID_GATE_MATRIX;
    }
    return launch_pair(st, bitpos(st->n, qubit), 0, m, kOpMatrix);
}

int lq_apply_ctrl_1q(lq_state* st, const int* controls, int num_controls, int target, const double m[8]) {
    LQ_API_GUARD;
    clear_error();
    int rc = check_state(st);
    if (rc) return rc;
    rc = check_qubit(st, target);
    if (rc) return rc;
    if (!m || num_controls < 0 || num_controls > 3 || (num_controls > 0 && !controls)) {
        set_error("invalid controls or matrix");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    uint64_t cm = 0;
    for (int i = 0; i < num_controls; ++i) {
        rc = check_qubit(st, controls[i]);
        if (rc) return rc;
        if (controls[i] == target || (cm & bitmask(st->n, controls[i]))) {
            set_error("Control and target must be different");
            return LQ_ERR_INVALID_ARGUMENT;
        }
        cm |= bitmask(st->n, controls[i]);
    }
    return launch_pair(st, bitpos(st->n, target), cm, m, kOpMatrix);
}

int lq_apply_2q(lq_state* st, int q_a, int q_b, const double m[32]) {
    LQ_API_GUARD;
    clear_error();
    int rc = check_state(st);
    if (rc) return rc;
    rc = check_qubit(st, q_a);
    if (rc) return rc;
    rc = check_qubit(st, q_b);
    if (rc) return rc;
    if (q_a == q_b || !m) {
        set_error("Qubits must be different");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    return launch_quad(st, bitpos(st->n, q_a), bitpos(st->n, q_b), m);
}

int lq_apply_dense(lq_state* st, const double* matrix, uint64_t dim) {
    LQ_API_GUARD;
    clear_error();
    int rc = check_state(st);
    if (rc) return rc;
    if (!matrix) {
        set_error("matrix is null");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    if (dim != st->dim) {
        // MatrixGate panics "Matrix columns (..) must match state dimension (..)"
        // (src/gates/mod.rs:81-87); apply_full_matrix returns DimensionMismatch
        set_error("Matrix columns (%llu) must match state dimension (%llu)", (unsigned long long)dim,
                  (unsigned long long)st->dim);
        return LQ_ERR_DIMENSION_MISMATCH;
    }
    if (st->n > 14) {
        set_error("dense matrices are limited to 14 qubits (a 2^n x 2^n complex matrix)");
        return LQ_ERR_UNSUPPORTED;
    }
    size_t bytes = (size_t)dim * dim * 16;
    void* dm = nullptr;
    LQ_HIP_CHECK(hipMalloc(&dm, bytes));
    hipError_t e = hipMemcpyAsync(dm, matrix, bytes, hipMemcpyHostToDevice, stream());
    if (e == hipSuccess) rc = launch_dense(st, (const double*)dm, dim);
    else {
        set_error("matrix upload failed: %s", hipGetErrorString(e));
        rc = LQ_ERR_DEVICE;
    }
    (void)hipStreamSynchronize(stream());
    (void)hipFree(dm);
    return rc;
}

int lq_execute(lq_state* st, const lq_gate* gates, int count, int mode) {
    LQ_API_GUARD;
    clear_error();
    int rc = check_state(st);
    if (rc) return rc;
    if (count < 0 || (count > 0 && !gates)) {
        set_error("invalid gate list");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    for (int i = 0; i < count; ++i) {
        rc = validate_gate(st->n, &gates[i]);
        if (rc) return rc;
    }
    if (mode == LQ_EXEC_PER_GATE) {
        std::vector<Prim> prims;
        rc = gates_to_prims(st->n, gates, count, prims);
        if (rc) return rc;
        return run_prims_per_gate(st, prims);
    }
    if (mode == LQ_EXEC_FUSED) return execute_fused(st, gates, count);
    set_error("unknown execution mode %d", mode);
    return LQ_ERR_INVALID_ARGUMENT;
}

}  // extern "C"
