// C ABI for gates and circuits: validates descriptors the way the reference's
// builders do and dispatches each gate kind to its kernel.
#include <cmath>
#include <cstring>
#include <vector>

#include "lq_common.h"
#include "lq_host.h"

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

// 2x2 matrix of a single-qubit gate kind (host, f64, reference formulas).
static void mat_rx(double th, double* m) {
    double c = std::cos(th / 2.0), s = std::sin(th / 2.0);
    double r[8] = {c, 0, 0, -s, 0, -s, c, 0};
    std::memcpy(m, r, sizeof(r));
}
static void mat_ry(double th, double* m) {
    double c = std::cos(th / 2.0), s = std::sin(th / 2.0);
    double r[8] = {c, 0, -s, 0, s, 0, c, 0};
    std::memcpy(m, r, sizeof(r));
}
static void mat_u3(double th, double ph, double la, double* m) {
    // U3Gate (parameterized.rs:254-261)
    double c = std::cos(th / 2.0), s = std::sin(th / 2.0);
    double p[2];
    m[0] = c;
    m[1] = 0.0;
    polar(s, la, p);
    m[2] = -p[0];
    m[3] = -p[1];
    polar(s, ph, p);
    m[4] = p[0];
    m[5] = p[1];
    polar(c, ph + la, p);
    m[6] = p[0];
    m[7] = p[1];
}

// Per-gate dispatch.  Angles are taken from g->theta.
int apply_gate_desc(lq_state* st, const lq_gate* g) {
    const int n = st->n;
    int rc = validate_gate(n, g);
    if (rc) return rc;
    const int* q = g->qubits;
    const double* th = g->theta;
    double m[8];
    double d[8] = {1, 0, 1, 0, 1, 0, 1, 0};
    switch (g->kind) {
        case LQ_GATE_X:
            rc = launch_pair(st, bitpos(n, q[0]), 0, m, kOpSwap);
            break;
        case LQ_GATE_Y:
            rc = launch_pair(st, bitpos(n, q[0]), 0, m, kOpNegY);
            break;
        case LQ_GATE_Z:
            d[0] = -1.0;
            d[1] = 0.0;
            rc = launch_diag(st, bitmask(n, q[0]), -1, -1, d);
            break;
        case LQ_GATE_H: {
            double s = 1.0 / std::sqrt(2.0);  // 1.0 / SQRT_2 (single_qubit.rs:237)
            double hm[8] = {s, 0, s, 0, s, 0, -s, 0};
            rc = launch_pair(st, bitpos(n, q[0]), 0, hm, 3 /* Hadamard op */);
            break;
        }
        case LQ_GATE_S:
            d[0] = 0.0;
            d[1] = 1.0;
            rc = launch_diag(st, bitmask(n, q[0]), -1, -1, d);
            break;
        case LQ_GATE_T:
            polar(1.0, M_PI / 4.0, d);
            rc = launch_diag(st, bitmask(n, q[0]), -1, -1, d);
            break;
        case LQ_GATE_RX:
            mat_rx(th[0], m);
            rc = launch_pair(st, bitpos(n, q[0]), 0, m, kOpMatrix);
            break;
        case LQ_GATE_RY:
            mat_ry(th[0], m);
            rc = launch_pair(st, bitpos(n, q[0]), 0, m, kOpMatrix);
            break;
        case LQ_GATE_RZ:
            polar(1.0, -th[0] / 2.0, d);
            polar(1.0, th[0] / 2.0, d + 2);
            rc = launch_diag(st, 0, bitpos(n, q[0]), -1, d);
            break;
        case LQ_GATE_PHASE:
            polar(1.0, th[0], d);
            rc = launch_diag(st, bitmask(n, q[0]), -1, -1, d);
            break;
        case LQ_GATE_U3:
            mat_u3(th[0], th[1], th[2], m);
            rc = launch_pair(st, bitpos(n, q[0]), 0, m, kOpMatrix);
            break;
        case LQ_GATE_CNOT:
            rc = launch_pair(st, bitpos(n, q[1]), bitmask(n, q[0]), m, kOpSwap);
            break;
        case LQ_GATE_SWAP:
            rc = launch_swap2(st, bitpos(n, q[0]), bitpos(n, q[1]));
            break;
        case LQ_GATE_CZ:
            d[0] = -1.0;
            d[1] = 0.0;
            rc = launch_diag(st, bitmask(n, q[0]) | bitmask(n, q[1]), -1, -1, d);
            break;
        case LQ_GATE_CPHASE:
        case LQ_GATE_CP:
            polar(1.0, th[0], d);
            rc = launch_diag(st, bitmask(n, q[0]) | bitmask(n, q[1]), -1, -1, d);
            break;
        case LQ_GATE_CRX:
            mat_rx(th[0], m);
            rc = launch_pair(st, bitpos(n, q[1]), bitmask(n, q[0]), m, kOpMatrix);
            break;
        case LQ_GATE_CRY:
            mat_ry(th[0], m);
            rc = launch_pair(st, bitpos(n, q[1]), bitmask(n, q[0]), m, kOpMatrix);
            break;
        case LQ_GATE_CRZ:
            polar(1.0, -th[0] / 2.0, d);
            polar(1.0, th[0] / 2.0, d + 2);
            rc = launch_diag(st, bitmask(n, q[0]), bitpos(n, q[1]), -1, d);
            break;
        case LQ_GATE_RXX:
        case LQ_GATE_RYY: {
            double M[32];
            two_qubit_rot_matrix(g->kind, th[0], q[0], q[1], M);
            rc = launch_quad(st, bitpos(n, q[0]), bitpos(n, q[1]), M);
            break;
        }
        case LQ_GATE_RZZ:
            // same parity (sub-index 0 or 3): e^{-i th/2}; different: e^{+i th/2}
            polar(1.0, -th[0] / 2.0, d);
            polar(1.0, th[0] / 2.0, d + 2);
            polar(1.0, th[0] / 2.0, d + 4);
            polar(1.0, -th[0] / 2.0, d + 6);
            rc = launch_diag(st, 0, bitpos(n, q[0]), bitpos(n, q[1]), d);
            break;
        case LQ_GATE_TOFFOLI:
            rc = launch_pair(st, bitpos(n, q[2]), bitmask(n, q[0]) | bitmask(n, q[1]), m, kOpSwap);
            break;
        case LQ_GATE_MATRIX1:
            rc = launch_pair(st, bitpos(n, q[0]), 0, g->matrix, kOpMatrix);
            if (!rc) rc = normalize_device(st, nullptr);
            break;
        case LQ_GATE_MATRIX2: {
            int q1 = q[0] < q[1] ? q[0] : q[1], q2 = q[0] < q[1] ? q[1] : q[0];
            rc = launch_quad(st, bitpos(n, q1), bitpos(n, q2), g->matrix);
            if (!rc) rc = normalize_device(st, nullptr);
            break;
        }
        default:
            set_error("unknown gate kind %d", g->kind);
            return LQ_ERR_INVALID_ARGUMENT;
    }
    if (rc) return rc;
    if ((g->flags & LQ_GATE_FLAG_RENORMALIZE) && g->kind != LQ_GATE_MATRIX1 && g->kind != LQ_GATE_MATRIX2)
        rc = normalize_device(st, nullptr);
    return rc;
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
        for (int i = 0; i < count; ++i) {
            rc = apply_gate_desc(st, &gates[i]);
            if (rc) return rc;
        }
        return LQ_OK;
    }
    if (mode == LQ_EXEC_FUSED) return execute_fused(st, gates, count);
    set_error("unknown execution mode %d", mode);
    return LQ_ERR_INVALID_ARGUMENT;
}

}  // extern "C"
