/*
 * lq_oracle.c -- CPU restatement of LogosQ 0.2.5's hot path (test
 * infrastructure only; see lq_oracle.h).  Built with -ffp-contract=off.
 *
 * Each loop keeps the reference's iteration order and arithmetic.  Two
 * orders the reference leaves to its libraries are fixed here and are
 * last-ulp effects only: rayon's reduction tree in State::normalize
 * (src/states/mod.rs:94-99) is a sequential sum, and ndarray's `dot`
 * (src/gates/mod.rs:89) sums each row in column order.
 */
#include "lq_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ---- num-complex arithmetic ------------------------------------------- */
static inline orc C(double re, double im) {
    orc z = {re, im};
    return z;
}
static inline orc cmul(orc a, orc b) { return C(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline orc cadd(orc a, orc b) { return C(a.re + b.re, a.im + b.im); }
static inline orc csub(orc a, orc b) { return C(a.re - b.re, a.im - b.im); }
static inline orc cneg(orc a) { return C(-a.re, -a.im); }
static inline orc cconj(orc a) { return C(a.re, -a.im); }
static inline orc cdivr(orc a, double r) { return C(a.re / r, a.im / r); }
static inline orc cpolar(double r, double th) { return C(r * cos(th), r * sin(th)); }

/* ---- State ---------------------------------------------------------------- */
double or_normalize(orc* v, uint64_t dim) {
    /* src/states/mod.rs:91-114: norm = sqrt(sum |c|^2); divide if > 1e-10 */
    double s = 0.0;
    for (uint64_t i = 0; i < dim; ++i) s += v[i].re * v[i].re + v[i].im * v[i].im;
    double norm = sqrt(s);
    if (norm > 1e-10)
        for (uint64_t i = 0; i < dim; ++i) v[i] = cdivr(v[i], norm);
    return norm;
}

void or_zero_state(orc* v, int n) {
    uint64_t dim = 1ull << n;
    memset(v, 0, dim * sizeof(orc));
    v[0] = C(1.0, 0.0);
}

void or_one_state(orc* v, int n) {
    uint64_t dim = 1ull << n;
    memset(v, 0, dim * sizeof(orc));
    v[dim - 1] = C(1.0, 0.0);
}

void or_plus_state(orc* v, int n) {
    uint64_t dim = 1ull << n;
    double a = 1.0 / sqrt((double)dim); /* mod.rs:154 */
    for (uint64_t i = 0; i < dim; ++i) v[i] = C(a, 0.0);
}

void or_inner_product(const orc* a, const orc* b, uint64_t dim, double out[2]) {
    orc s = C(0.0, 0.0);
    for (uint64_t i = 0; i < dim; ++i) s = cadd(s, cmul(cconj(a[i]), b[i]));
    out[0] = s.re;
    out[1] = s.im;
}

/* ---- single-qubit gates ---------------------------------------------------- */
void or_x(orc* v, int n, int q) {
    uint64_t stride = 1ull << (n - 1 - q), dim = 1ull << n;
    for (uint64_t base = 0; base < dim; base += 2 * stride)
        for (uint64_t off = 0; off < stride; ++off) {
            uint64_t i0 = base + off, i1 = i0 + stride;
            orc t = v[i0];
            v[i0] = v[i1];
            v[i1] = t;
        }
}

void or_y(orc* v, int n, int q) {
    uint64_t stride = 1ull << (n - 1 - q), dim = 1ull << n;
    for (uint64_t base = 0; base < dim; base += 2 * stride)
        for (uint64_t off = 0; off < stride; ++off) {
            uint64_t i0 = base + off, i1 = i0 + stride;
            orc t = v[i0];
            v[i0] = cmul(C(0.0, 1.0), v[i1]);
            v[i1] = cmul(C(0.0, -1.0), t);
        }
}

void or_z(orc* v, int n, int q) {
    uint64_t mask = 1ull << (n - 1 - q), dim = 1ull << n;
    for (uint64_t i = 0; i < dim; ++i)
        if (i & mask) v[i] = cneg(v[i]);
}

void or_h(orc* v, int n, int q) {
    uint64_t stride = 1ull << (n - 1 - q), dim = 1ull << n;
    orc s = C(1.0 / sqrt(2.0), 0.0); /* 1.0 / SQRT_2, single_qubit.rs:237 */
    for (uint64_t base = 0; base < dim; base += 2 * stride)
        for (uint64_t off = 0; off < stride; ++off) {
            uint64_t i0 = base + off, i1 = i0 + stride;
            orc a0 = v[i0], a1 = v[i1];
            v[i0] = cmul(s, cadd(a0, a1));
            v[i1] = cmul(s, csub(a0, a1));
        }
}

static void diag_bit1(orc* v, int n, int q, orc ph) {
    uint64_t mask = 1ull << (n - 1 - q), dim = 1ull << n;
    for (uint64_t i = 0; i < dim; ++i)
        if (i & mask) v[i] = cmul(v[i], ph);
}

void or_s(orc* v, int n, int q) { diag_bit1(v, n, q, C(0.0, 1.0)); }
void or_t(orc* v, int n, int q) { diag_bit1(v, n, q, cpolar(1.0, M_PI / 4.0)); }
void or_phase(orc* v, int n, int q, double theta) { diag_bit1(v, n, q, cpolar(1.0, theta)); }

static void pair_update(orc* v, int n, int q, orc m00, orc m01, orc m10, orc m11) {
    uint64_t stride = 1ull << (n - 1 - q), dim = 1ull << n;
    for (uint64_t base = 0; base < dim; base += 2 * stride)
        for (uint64_t off = 0; off < stride; ++off) {
            uint64_t i0 = base + off, i1 = i0 + stride;
            orc a0 = v[i0], a1 = v[i1];
            v[i0] = cadd(cmul(m00, a0), cmul(m01, a1));
            v[i1] = cadd(cmul(m10, a0), cmul(m11, a1));
        }
}

void or_rx(orc* v, int n, int q, double theta) {
    double c = cos(theta / 2.0), s = sin(theta / 2.0);
    pair_update(v, n, q, C(c, 0.0), C(0.0, -s), C(0.0, -s), C(c, 0.0));
}

void or_ry(orc* v, int n, int q, double theta) {
    double c = cos(theta / 2.0), s = sin(theta / 2.0);
    pair_update(v, n, q, C(c, 0.0), C(-s, 0.0), C(s, 0.0), C(c, 0.0));
}

void or_rz(orc* v, int n, int q, double theta) {
    orc p0 = cpolar(1.0, -theta / 2.0), p1 = cpolar(1.0, theta / 2.0);
    uint64_t mask = 1ull << (n - 1 - q), dim = 1ull << n;
    for (uint64_t i = 0; i < dim; ++i) v[i] = (i & mask) ? cmul(v[i], p1) : cmul(v[i], p0);
}

void or_u3(orc* v, int n, int q, double theta, double phi, double lam) {
    double c = cos(theta / 2.0), s = sin(theta / 2.0);
    orc m00 = C(c, 0.0);
    orc m01 = cneg(cpolar(s, lam));
    orc m10 = cpolar(s, phi);
    orc m11 = cpolar(c, phi + lam);
    pair_update(v, n, q, m00, m01, m10, m11);
}

/* ---- two- and three-qubit gates ------------------------------------------- */
void or_cnot(orc* v, int n, int c, int t) {
    uint64_t cm = 1ull << (n - 1 - c), tm = 1ull << (n - 1 - t), dim = 1ull << n;
    for (uint64_t i = 0; i < dim; ++i)
        if ((i & cm) && !(i & tm)) {
            uint64_t j = i ^ tm;
            orc x = v[i];
            v[i] = v[j];
            v[j] = x;
        }
}

void or_swap(orc* v, int n, int q1, int q2) {
    uint64_t m1 = 1ull << (n - 1 - q1), m2 = 1ull << (n - 1 - q2), dim = 1ull << n;
    for (uint64_t i = 0; i < dim; ++i)
        if ((i & m1) && !(i & m2)) {
            uint64_t j = i ^ m1 ^ m2;
            orc x = v[i];
            v[i] = v[j];
            v[j] = x;
        }
}

void or_cz(orc* v, int n, int c, int t) {
    uint64_t both = (1ull << (n - 1 - c)) | (1ull << (n - 1 - t)), dim = 1ull << n;
    for (uint64_t i = 0; i < dim; ++i)
        if ((i & both) == both) v[i] = cneg(v[i]);
}

void or_cphase(orc* v, int n, int c, int t, double angle) {
    /* the rayon branch (N > 1024) multiplies the same elements by the same
     * phase; element-wise, so the order does not matter */
    uint64_t both = (1ull << (n - 1 - c)) | (1ull << (n - 1 - t)), dim = 1ull << n;
    orc ph = cpolar(1.0, angle);
    for (uint64_t i = 0; i < dim; ++i)
        if ((i & both) == both) v[i] = cmul(v[i], ph);
}

static void ctrl_pair(orc* v, int n, int c, int t, orc m00, orc m01, orc m10, orc m11) {
    uint64_t cm = 1ull << (n - 1 - c), tm = 1ull << (n - 1 - t), dim = 1ull << n;
    for (uint64_t i = 0; i < dim; ++i)
        if ((i & cm) && !(i & tm)) {
            uint64_t j = i | tm;
            orc a0 = v[i], a1 = v[j];
            v[i] = cadd(cmul(m00, a0), cmul(m01, a1));
            v[j] = cadd(cmul(m10, a0), cmul(m11, a1));
        }
}

void or_crx(orc* v, int n, int c, int t, double theta) {
    double co = cos(theta / 2.0), s = sin(theta / 2.0);
    ctrl_pair(v, n, c, t, C(co, 0.0), C(0.0, -s), C(0.0, -s), C(co, 0.0));
}

void or_cry(orc* v, int n, int c, int t, double theta) {
    double co = cos(theta / 2.0), s = sin(theta / 2.0);
    ctrl_pair(v, n, c, t, C(co, 0.0), C(-s, 0.0), C(s, 0.0), C(co, 0.0));
}

void or_crz(orc* v, int n, int c, int t, double theta) {
    uint64_t cm = 1ull << (n - 1 - c), tm = 1ull << (n - 1 - t), dim = 1ull << n;
    orc p0 = cpolar(1.0, -theta / 2.0), p1 = cpolar(1.0, theta / 2.0);
    for (uint64_t i = 0; i < dim; ++i)
        if (i & cm) v[i] = (i & tm) ? cmul(v[i], p1) : cmul(v[i], p0);
}

void or_cp(orc* v, int n, int c, int t, double theta) {
    uint64_t both = (1ull << (n - 1 - c)) | (1ull << (n - 1 - t)), dim = 1ull << n;
    orc ph = cpolar(1.0, theta);
    for (uint64_t i = 0; i < dim; ++i)
        if ((i & both) == both) v[i] = cmul(v[i], ph);
}

/* RXX / RYY: the 00<->11 pair always; the 10<->01 pair only where i < j,
 * i.e. only when qubit1 > qubit2 (parameterized.rs:563-575, 638-650). */
static void two_block(orc* v, int n, int q1, int q2, orc d, orc o_0011, orc o_1001) {
    uint64_t m1 = 1ull << (n - 1 - q1), m2 = 1ull << (n - 1 - q2), dim = 1ull << n;
    for (uint64_t i = 0; i < dim; ++i) {
        int b1 = (i & m1) != 0, b2 = (i & m2) != 0;
        if (!b1 && !b2) {
            uint64_t j = i | m1 | m2;
            if (i < j) {
                orc ai = v[i], aj = v[j];
                v[i] = cadd(cmul(d, ai), cmul(o_0011, aj));
                v[j] = cadd(cmul(o_0011, ai), cmul(d, aj));
            }
        } else if (b1 && !b2) {
            uint64_t j = (i & ~m1) | m2;
            if (i < j) {
                orc ai = v[i], aj = v[j];
                v[i] = cadd(cmul(d, ai), cmul(o_1001, aj));
                v[j] = cadd(cmul(o_1001, ai), cmul(d, aj));
            }
        }
    }
}

void or_rxx(orc* v, int n, int q1, int q2, double theta) {
    double c = cos(theta / 2.0), s = sin(theta / 2.0);
    two_block(v, n, q1, q2, C(c, 0.0), C(0.0, -s), C(0.0, -s));
}

void or_ryy(orc* v, int n, int q1, int q2, double theta) {
    double c = cos(theta / 2.0), s = sin(theta / 2.0);
    two_block(v, n, q1, q2, C(c, 0.0), C(0.0, s), C(0.0, -s));
}

void or_rzz(orc* v, int n, int q1, int q2, double theta) {
    uint64_t m1 = 1ull << (n - 1 - q1), m2 = 1ull << (n - 1 - q2), dim = 1ull << n;
    for (uint64_t i = 0; i < dim; ++i) {
        int b1 = (i & m1) != 0, b2 = (i & m2) != 0;
        /* from_polar is evaluated per element in the reference loop; the
         * value is the same for every element */
        orc ph = (b1 == b2) ? cpolar(1.0, -theta / 2.0) : cpolar(1.0, theta / 2.0);
        v[i] = cmul(v[i], ph);
    }
}

void or_toffoli(orc* v, int n, int c1, int c2, int t) {
    uint64_t cm = (1ull << (n - 1 - c1)) | (1ull << (n - 1 - c2)), tm = 1ull << (n - 1 - t), dim = 1ull << n;
    for (uint64_t i = 0; i < dim; ++i)
        if ((i & cm) == cm && !(i & tm)) {
            uint64_t j = i ^ tm;
            orc x = v[i];
            v[i] = v[j];
            v[j] = x;
        }
}

/* ---- dense matrices --------------------------------------------------------- */
void or_matrix_gate(orc* v, int n, const orc* m, uint64_t dim, orc* scratch) {
    /* src/gates/mod.rs:77-93 */
    for (uint64_t i = 0; i < dim; ++i) {
        orc s = C(0.0, 0.0);
        const orc* row = m + i * dim;
        for (uint64_t j = 0; j < dim; ++j) s = cadd(s, cmul(row[j], v[j]));
        scratch[i] = s;
    }
    memcpy(v, scratch, dim * sizeof(orc));
    or_normalize(v, dim);
    (void)n;
}

/* tensor_product (src/circuits/utils.rs:11-49): result[i*rb+k][j*cb+l] = a[i][j]*b[k][l] */
static void kron(const orc* a, int ar, int ac, const orc* b, int br, int bc, orc* out) {
    int rr = ar * br, rc = ac * bc;
    for (int i = 0; i < ar; ++i)
        for (int j = 0; j < ac; ++j)
            for (int k = 0; k < br; ++k)
                for (int l = 0; l < bc; ++l) out[(size_t)(i * br + k) * rc + (j * bc + l)] = cmul(a[i * ac + j], b[k * bc + l]);
    (void)rr;
}

void or_kron_single(const orc m[4], int n, int target, orc* full) {
    /* src/circuits/single_qubit.rs:28-58 */
    orc id[4] = {C(1, 0), C(0, 0), C(0, 0), C(1, 0)};
    uint64_t dim = 1ull << n;
    orc* cur = (orc*)malloc(dim * dim * sizeof(orc));
    orc* nxt = (orc*)malloc(dim * dim * sizeof(orc));
    int sz = 2;
    memcpy(cur, target == 0 ? m : id, 4 * sizeof(orc));
    for (int i = 1; i < n; ++i) {
        kron(cur, sz, sz, i == target ? m : id, 2, 2, nxt);
        sz *= 2;
        orc* t = cur;
        cur = nxt;
        nxt = t;
    }
    memcpy(full, cur, dim * dim * sizeof(orc));
    free(cur);
    free(nxt);
}

void or_expand_two(const orc m[16], int n, int qa, int qb, orc* full) {
    /* src/circuits/two_qubits.rs:20-79: sorted pair, column i, rows from m[j][subidx] */
    int q1 = qa < qb ? qa : qb, q2 = qa < qb ? qb : qa;
    uint64_t dim = 1ull << n;
    if (n == 2) {
        memcpy(full, m, 16 * sizeof(orc));
        return;
    }
    memset(full, 0, dim * dim * sizeof(orc));
    int bp1 = n - 1 - q1, bp2 = n - 1 - q2;
    for (uint64_t i = 0; i < dim; ++i) {
        uint64_t b1 = (i >> bp1) & 1, b2 = (i >> bp2) & 1;
        uint64_t sub = (b1 << 1) | b2;
        for (uint64_t j = 0; j < 4; ++j) {
            orc e = m[j * 4 + sub];
            if (e.re == 0.0 && e.im == 0.0) continue;
            uint64_t nb1 = (j >> 1) & 1, nb2 = j & 1;
            uint64_t ns = i;
            ns = (ns & ~(1ull << bp1)) | (nb1 << bp1);
            ns = (ns & ~(1ull << bp2)) | (nb2 << bp2);
            full[ns * dim + i] = e;
        }
    }
}

void or_matrix1(orc* v, int n, int q, const orc m[4]) {
    pair_update(v, n, q, m[0], m[1], m[2], m[3]);
    or_normalize(v, 1ull << n);
}

void or_matrix2(orc* v, int n, int qa, int qb, const orc m[16]) {
    int q1 = qa < qb ? qa : qb, q2 = qa < qb ? qb : qa;
    uint64_t m1 = 1ull << (n - 1 - q1), m2 = 1ull << (n - 1 - q2), dim = 1ull << n;
    for (uint64_t i = 0; i < dim; ++i) {
        if (i & (m1 | m2)) continue;
        uint64_t idx[4] = {i, i | m2, i | m1, i | m1 | m2};
        orc a[4];
        for (int k = 0; k < 4; ++k) a[k] = v[idx[k]];
        for (int r = 0; r < 4; ++r) {
            /* dot over columns in ascending index order; only the four
             * columns of this block are non-zero */
            orc s = C(0.0, 0.0);
            for (int k = 0; k < 4; ++k) s = cadd(s, cmul(m[r * 4 + k], a[k]));
            v[idx[r]] = s;
        }
    }
    or_normalize(v, dim);
}

void or_matrix_of(int which, double theta, orc m[4]) {
    /* src/gates/matrix.rs:14-161 */
    double r2 = 1.0 / sqrt(2.0);
    double c = cos(theta / 2.0), s = sin(theta / 2.0);
    switch (which) {
        case LQ_GATE_X: m[0] = C(0, 0); m[1] = C(1, 0); m[2] = C(1, 0); m[3] = C(0, 0); break;
        case LQ_GATE_Y: m[0] = C(0, 0); m[1] = C(0, -1); m[2] = C(0, 1); m[3] = C(0, 0); break;
        case LQ_GATE_Z: m[0] = C(1, 0); m[1] = C(0, 0); m[2] = C(0, 0); m[3] = C(-1, 0); break;
        case LQ_GATE_H: m[0] = C(r2, 0); m[1] = C(r2, 0); m[2] = C(r2, 0); m[3] = C(-r2, 0); break;
        case LQ_GATE_S: m[0] = C(1, 0); m[1] = C(0, 0); m[2] = C(0, 0); m[3] = C(0, 1); break;
        case LQ_GATE_T: m[0] = C(1, 0); m[1] = C(0, 0); m[2] = C(0, 0); m[3] = C(r2, r2); break;
        case LQ_GATE_RX: m[0] = C(c, 0); m[1] = C(0, -s); m[2] = C(0, -s); m[3] = C(c, 0); break;
        case LQ_GATE_RY: m[0] = C(c, 0); m[1] = C(-s, 0); m[2] = C(s, 0); m[3] = C(c, 0); break;
        case LQ_GATE_RZ:
            m[0] = C(cos(theta / 2.0), -sin(theta / 2.0));
            m[1] = C(0, 0);
            m[2] = C(0, 0);
            m[3] = C(cos(theta / 2.0), sin(theta / 2.0));
            break;
        default: m[0] = C(1, 0); m[1] = C(0, 0); m[2] = C(0, 0); m[3] = C(1, 0); break;
    }
}

/* ---- circuits -------------------------------------------------------------- */
int or_apply_gate(orc* v, int n, const lq_gate* g) {
    const int* q = g->qubits;
    const double* th = g->theta;
    switch (g->kind) {
        case LQ_GATE_X: or_x(v, n, q[0]); break;
        case LQ_GATE_Y: or_y(v, n, q[0]); break;
        case LQ_GATE_Z: or_z(v, n, q[0]); break;
        case LQ_GATE_H: or_h(v, n, q[0]); break;
        case LQ_GATE_S: or_s(v, n, q[0]); break;
        case LQ_GATE_T: or_t(v, n, q[0]); break;
        case LQ_GATE_RX: or_rx(v, n, q[0], th[0]); break;
        case LQ_GATE_RY: or_ry(v, n, q[0], th[0]); break;
        case LQ_GATE_RZ: or_rz(v, n, q[0], th[0]); break;
        case LQ_GATE_PHASE: or_phase(v, n, q[0], th[0]); break;
        case LQ_GATE_U3: or_u3(v, n, q[0], th[0], th[1], th[2]); break;
        case LQ_GATE_CNOT: or_cnot(v, n, q[0], q[1]); break;
        case LQ_GATE_SWAP: or_swap(v, n, q[0], q[1]); break;
        case LQ_GATE_CZ: or_cz(v, n, q[0], q[1]); break;
        case LQ_GATE_CPHASE: or_cphase(v, n, q[0], q[1], th[0]); break;
        case LQ_GATE_CRX: or_crx(v, n, q[0], q[1], th[0]); break;
        case LQ_GATE_CRY: or_cry(v, n, q[0], q[1], th[0]); break;
        case LQ_GATE_CRZ: or_crz(v, n, q[0], q[1], th[0]); break;
        case LQ_GATE_CP: or_cp(v, n, q[0], q[1], th[0]); break;
        case LQ_GATE_RXX: or_rxx(v, n, q[0], q[1], th[0]); break;
        case LQ_GATE_RYY: or_ryy(v, n, q[0], q[1], th[0]); break;
        case LQ_GATE_RZZ: or_rzz(v, n, q[0], q[1], th[0]); break;
        case LQ_GATE_TOFFOLI: or_toffoli(v, n, q[0], q[1], q[2]); break;
        case LQ_GATE_MATRIX1: or_matrix1(v, n, q[0], (const orc*)g->matrix); break;
        case LQ_GATE_MATRIX2: or_matrix2(v, n, q[0], q[1], (const orc*)g->matrix); break;
        default: return -1;
    }
    if ((g->flags & LQ_GATE_FLAG_RENORMALIZE) && g->kind != LQ_GATE_MATRIX1 && g->kind != LQ_GATE_MATRIX2)
        or_normalize(v, 1ull << n);
    return 0;
}

int or_execute(orc* v, int n, const lq_gate* gates, int count) {
    for (int i = 0; i < count; ++i)
        if (or_apply_gate(v, n, &gates[i]) != 0) return -1;
    return 0;
}

static void set_gate(lq_gate* g, int kind, int q0, int q1, double th) {
    memset(g, 0, sizeof(*g));
    g->kind = kind;
    g->qubits[0] = q0;
    g->qubits[1] = q1;
    g->qubits[2] = -1;
    g->theta[0] = th;
    g->param[0] = g->param[1] = g->param[2] = -1;
}

static void set_dense_h(lq_gate* g, int q) {
    orc m[4];
    set_gate(g, LQ_GATE_MATRIX1, q, -1, 0.0);
    or_matrix_of(LQ_GATE_H, 0.0, m);
    memcpy(g->matrix, m, sizeof(m));
    g->flags = LQ_GATE_FLAG_RENORMALIZE;
}

int or_qft_circuit(int n, lq_gate* gates) {
    /* src/algorithms/qft.rs:10-27: H(i) is the dense builder Circuit::h;
     * controlled phases are CPhaseGate(i, j, pi / 2^(j-i)); then swaps. */
    int k = 0;
    for (int i = 0; i < n; ++i) {
        if (gates) set_dense_h(&gates[k], i);
        ++k;
        for (int j = i + 1; j < n; ++j) {
            double angle = M_PI / (double)(1ull << (j - i));
            if (gates) set_gate(&gates[k], LQ_GATE_CPHASE, i, j, angle);
            ++k;
        }
    }
    for (int i = 0; i < n / 2; ++i) {
        if (gates) set_gate(&gates[k], LQ_GATE_SWAP, i, n - i - 1, 0.0);
        ++k;
    }
    return k;
}

int or_inverse_qft_circuit(int n, lq_gate* gates) {
    /* src/algorithms/qft.rs:30-46 */
    int k = 0;
    for (int i = 0; i < n / 2; ++i) {
        if (gates) set_gate(&gates[k], LQ_GATE_SWAP, i, n - i - 1, 0.0);
        ++k;
    }
    for (int i = n - 1; i >= 0; --i) {
        for (int j = i + 1; j < n; ++j) {
            double angle = -M_PI / (double)(1ull << (j - i));
            if (gates) set_gate(&gates[k], LQ_GATE_CPHASE, i, j, angle);
            ++k;
        }
        if (gates) set_dense_h(&gates[k], i);
        ++k;
    }
    return k;
}

/* ---- observables ---------------------------------------------------------- */
double or_pauli_term_expectation(const orc* v, int n, const uint8_t* paulis, double coeff) {
    /* src/optimization/observable.rs:95-147 */
    uint64_t dim = 1ull << n;
    orc result = C(0.0, 0.0);
    for (uint64_t i = 0; i < dim; ++i) {
        orc amp = C(1.0, 0.0);
        uint64_t j = i;
        for (int q = 0; q < n; ++q) {
            uint64_t bit = (i >> (n - 1 - q)) & 1;
            switch (paulis[q]) {
                case LQ_PAULI_X: j ^= 1ull << (n - 1 - q); break;
                case LQ_PAULI_Y:
                    j ^= 1ull << (n - 1 - q);
                    amp = cmul(amp, bit == 0 ? C(0.0, 1.0) : C(0.0, -1.0));
                    break;
                case LQ_PAULI_Z:
                    if (bit == 1) amp = cmul(amp, C(-1.0, 0.0));
                    break;
                default: break;
            }
        }
        result = cadd(result, cmul(cmul(cconj(v[i]), amp), v[j]));
    }
    return coeff * result.re;
}

double or_pauli_sum_expectation(const orc* v, int n, const uint8_t* paulis, const double* coeffs, int num_terms) {
    double s = 0.0;
    for (int t = 0; t < num_terms; ++t) s += or_pauli_term_expectation(v, n, paulis + (size_t)t * n, coeffs[t]);
    return s;
}

/* ---- ansatz templates, gradients ------------------------------------------ */
static void bind(const lq_gate* tmpl, int num_gates, const double* params, lq_gate* out) {
    memcpy(out, tmpl, (size_t)num_gates * sizeof(lq_gate));
    for (int g = 0; g < num_gates; ++g)
        for (int k = 0; k < 3; ++k)
            if (out[g].param[k] >= 0) out[g].theta[k] = params[out[g].param[k]];
}

void or_run_template(orc* v, int n, const lq_gate* tmpl, int num_gates, const double* params, lq_gate* scratch) {
    or_zero_state(v, n);
    bind(tmpl, num_gates, params, scratch);
    or_execute(v, n, scratch, num_gates);
}

double or_energy(int n, const lq_gate* tmpl, int num_gates, const double* params, const uint8_t* paulis,
                 const double* coeffs, int num_terms) {
    orc* v = (orc*)malloc(sizeof(orc) << n);
    lq_gate* sc = (lq_gate*)malloc((size_t)num_gates * sizeof(lq_gate));
    or_run_template(v, n, tmpl, num_gates, params, sc);
    double e = or_pauli_sum_expectation(v, n, paulis, coeffs, num_terms);
    free(sc);
    free(v);
    return e;
}

void or_param_shift(int n, const lq_gate* tmpl, int num_gates, const double* params, int num_params, double shift,
                    const uint8_t* paulis, const double* coeffs, int num_terms, double* grad_out) {
    or_param_shift_range(n, tmpl, num_gates, params, num_params, shift, paulis, coeffs, num_terms, 0, num_params,
                         grad_out);
}

void or_param_shift_range(int n, const lq_gate* tmpl, int num_gates, const double* params, int num_params,
                          double shift, const uint8_t* paulis, const double* coeffs, int num_terms, int first,
                          int last, double* grad_out) {
    /* src/optimization/gradient.rs:45-78 (the loop over i, for i in [first, last)) */
    double* p = (double*)malloc((size_t)num_params * sizeof(double));
    for (int i = first; i < last; ++i) {
        memcpy(p, params, (size_t)num_params * sizeof(double));
        p[i] += shift;
        double ep = or_energy(n, tmpl, num_gates, p, paulis, coeffs, num_terms);
        memcpy(p, params, (size_t)num_params * sizeof(double));
        p[i] -= shift;
        double em = or_energy(n, tmpl, num_gates, p, paulis, coeffs, num_terms);
        grad_out[i] = (ep - em) / 2.0;
    }
    free(p);
}

void or_finite_difference(int n, const lq_gate* tmpl, int num_gates, const double* params, int num_params,
                          double eps, const uint8_t* paulis, const double* coeffs, int num_terms,
                          double* grad_out) {
    /* src/optimization/gradient.rs:99-125 */
    double fx = or_energy(n, tmpl, num_gates, params, paulis, coeffs, num_terms);
    double* p = (double*)malloc((size_t)num_params * sizeof(double));
    for (int i = 0; i < num_params; ++i) {
        memcpy(p, params, (size_t)num_params * sizeof(double));
        p[i] += eps;
        double fp = or_energy(n, tmpl, num_gates, p, paulis, coeffs, num_terms);
        grad_out[i] = (fp - fx) / eps;
    }
    free(p);
}

/* ---- optimizers ------------------------------------------------------------ */
/* f64::powi as Rust lowers it for a run-time exponent: binary
 * exponentiation by repeated squaring (LLVM's powi semantics). */
static double powi_f64(double a, int b) {
    int recip = b < 0;
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

or_opt_result or_gd_minimize(double lr, int max_iterations, double tol, or_objective_fn f, or_gradient_fn g,
                             void* user, double* x, int dim, double* history) {
    /* src/optimization/optimizer.rs:56-107 */
    or_opt_result r = {0.0, 0, 0};
    double* grad = (double*)malloc((size_t)dim * sizeof(double) + 8);
    int hist = 0;
    for (int it = 0; it < max_iterations; ++it) {
        double value = f(x, dim, user);
        if (history) history[hist] = value;
        ++hist;
        g(x, dim, grad, user);
        double gn = 0.0;
        for (int i = 0; i < dim; ++i) gn += grad[i] * grad[i];
        gn = sqrt(gn);
        if (gn < tol) {
            r.converged = 1;
            break;
        }
        for (int i = 0; i < dim; ++i) x[i] -= lr * grad[i];
    }
    r.optimal_value = f(x, dim, user);
    g(x, dim, grad, user);
    double fgn = 0.0;
    for (int i = 0; i < dim; ++i) fgn += grad[i] * grad[i];
    if (sqrt(fgn) < tol) r.converged = 1;
    r.num_iterations = hist;
    free(grad);
    return r;
}

or_opt_result or_adam_minimize(double lr, double beta1, double beta2, double eps, int max_iterations, double tol,
                               or_objective_fn f, or_gradient_fn g, void* user, double* x, int dim,
                               double* history) {
    /* src/optimization/optimizer.rs:139-223 */
    or_opt_result r = {0.0, 0, 0};
    double* m = (double*)calloc((size_t)dim + 1, sizeof(double));
    double* v = (double*)calloc((size_t)dim + 1, sizeof(double));
    double* grad = (double*)malloc((size_t)dim * sizeof(double) + 8);
    int hist = 0, have_prev = 0;
    double prev = 0.0;
    for (int it = 1; it <= max_iterations; ++it) {
        double value = f(x, dim, user);
        if (history) history[hist] = value;
        ++hist;
        if (have_prev && fabs(prev - value) < tol) {
            r.converged = 1;
            break;
        }
        prev = value;
        have_prev = 1;
        g(x, dim, grad, user);
        double gn = 0.0;
        for (int i = 0; i < dim; ++i) gn += grad[i] * grad[i];
        gn = sqrt(gn);
        if (gn < tol) {
            r.converged = 1;
            break;
        }
        for (int i = 0; i < dim; ++i) {
            m[i] = beta1 * m[i] + (1.0 - beta1) * grad[i];
            v[i] = beta2 * v[i] + (1.0 - beta2) * grad[i] * grad[i];
        }
        double mc = 1.0 / (1.0 - powi_f64(beta1, it));
        double vc = 1.0 / (1.0 - powi_f64(beta2, it));
        double sn = 0.0;
        for (int i = 0; i < dim; ++i) {
            double mh = m[i] * mc, vh = v[i] * vc;
            double step = lr * mh / (sqrt(vh) + eps);
            x[i] -= step;
            sn += step * step;
        }
        if (sqrt(sn) < tol) {
            r.converged = 1;
            break;
        }
    }
    r.optimal_value = f(x, dim, user);
    g(x, dim, grad, user);
    double fgn = 0.0;
    for (int i = 0; i < dim; ++i) fgn += grad[i] * grad[i];
    if (sqrt(fgn) < tol) r.converged = 1;
    r.num_iterations = hist;
    free(m);
    free(v);
    free(grad);
    return r;
}

typedef struct {
    int n, num_gates, num_terms;
    const lq_gate* tmpl;
    const uint8_t* paulis;
    const double* coeffs;
    double shift;
} vqe_ctx;

static double vqe_f(const double* x, int dim, void* user) {
    vqe_ctx* c = (vqe_ctx*)user;
    (void)dim;
    return or_energy(c->n, c->tmpl, c->num_gates, x, c->paulis, c->coeffs, c->num_terms);
}

static void vqe_g(const double* x, int dim, double* grad, void* user) {
    vqe_ctx* c = (vqe_ctx*)user;
    or_param_shift(c->n, c->tmpl, c->num_gates, x, dim, c->shift, c->paulis, c->coeffs, c->num_terms, grad);
}

or_opt_result or_vqe_run(int n, const lq_gate* tmpl, int num_gates, int num_params, const uint8_t* paulis,
                         const double* coeffs, int num_terms, double shift, int optimizer, double lr,
                         int max_iterations, double tol, double* params, double* history, orc* final_state) {
    /* src/optimization/vqe.rs:55-99 */
    vqe_ctx c = {n, num_gates, num_terms, tmpl, paulis, coeffs, shift};
    or_opt_result r;
    if (optimizer == 1)
        r = or_adam_minimize(lr, 0.9, 0.999, 1e-8, max_iterations, tol, vqe_f, vqe_g, &c, params, num_params,
                             history);
    else
        r = or_gd_minimize(lr, max_iterations, tol, vqe_f, vqe_g, &c, params, num_params, history);
    if (final_state) {
        lq_gate* sc = (lq_gate*)malloc((size_t)num_gates * sizeof(lq_gate));
        or_run_template(final_state, n, tmpl, num_gates, params, sc);
        free(sc);
    }
    return r;
}

/* ---- XYZ Heisenberg energy (src/algorithms/xyz_heisenberg.rs:143-252) ----- */
static double exp_xx(const orc* v, int n, int i, int j) {
    uint64_t mi = 1ull << (n - 1 - i), mj = 1ull << (n - 1 - j), dim = 1ull << n;
    orc r = C(0, 0);
    for (uint64_t k = 0; k < dim; ++k) r = cadd(r, cmul(cconj(v[k]), v[k ^ mi ^ mj]));
    return r.re;
}

static double exp_yy(const orc* v, int n, int i, int j) {
    uint64_t mi = 1ull << (n - 1 - i), mj = 1ull << (n - 1 - j), dim = 1ull << n;
    orc r = C(0, 0);
    for (uint64_t k = 0; k < dim; ++k) {
        orc pi = (k & mi) ? C(0, -1) : C(0, 1);
        orc pj = (k & mj) ? C(0, -1) : C(0, 1);
        orc ph = cmul(pi, pj);
        r = cadd(r, cmul(cmul(cconj(v[k]), ph), v[k ^ mi ^ mj]));
    }
    return r.re;
}

static double exp_zz(const orc* v, int n, int i, int j) {
    uint64_t mi = 1ull << (n - 1 - i), mj = 1ull << (n - 1 - j), dim = 1ull << n;
    double r = 0.0;
    for (uint64_t k = 0; k < dim; ++k) {
        int bi = (k & mi) != 0, bj = (k & mj) != 0;
        double sg = ((bi + bj) % 2 == 0) ? 1.0 : -1.0;
        r += (v[k].re * v[k].re + v[k].im * v[k].im) * sg;
    }
    return r;
}

static double exp_z(const orc* v, int n, int i) {
    uint64_t mi = 1ull << (n - 1 - i), dim = 1ull << n;
    double r = 0.0;
    for (uint64_t k = 0; k < dim; ++k) r += (v[k].re * v[k].re + v[k].im * v[k].im) * ((k & mi) ? -1.0 : 1.0);
    return r;
}

double or_xyz_energy_efficient(const orc* v, int n, double jx, double jy, double jz, double field) {
    double e = 0.0;
    for (int i = 0; i + 1 < n; ++i) {
        e += jx * exp_xx(v, n, i, i + 1);
        e += jy * exp_yy(v, n, i, i + 1);
        e += jz * exp_zz(v, n, i, i + 1);
    }
    if (field != 0.0)
        for (int i = 0; i < n; ++i) e += field * exp_z(v, n, i);
    return e;
}

double or_now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}
