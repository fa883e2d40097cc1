// Internal definitions shared by the HIP translation units of liblogosq_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>

#include "../../include/logosq_hip.h"

namespace lq {

// ---- amplitude types ---------------------------------------------------------
// complex128 amplitudes are double2 (16 B), complex64 are float2 (8 B); both
// interleaved (re, im) exactly like num_complex::Complex<T>.
template <typename T> struct cvec;
template <> struct cvec<double> { using type = double2; };
template <> struct cvec<float> { using type = float2; };
template <typename T> using cplx = typename cvec<T>::type;

// A 2x2 complex matrix passed by value (lands in scalar registers).
template <typename T> struct Mat2 {
    T m[8];  // row-major (m00, m01, m10, m11), interleaved re/im
};
template <typename T> struct Mat4 {
    T m[32];
};
// Four diagonal phases selected by two index bits.
template <typename T> struct Diag4 {
    T d[8];
};

__host__ __device__ inline double2 mk(double re, double im) { return make_double2(re, im); }
__host__ __device__ inline float2 mk(float re, float im) { return make_float2(re, im); }

// (a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re): num-complex's product.
template <typename V, typename T>
__device__ __forceinline__ V cmul(T ar, T ai, V b) {
    V r;
    r.x = ar * b.x - ai * b.y;
    r.y = ar * b.y + ai * b.x;
    return r;
}

template <typename V>
__device__ __forceinline__ V cadd(V a, V b) {
    V r;
    r.x = a.x + b.x;
    r.y = a.y + b.y;
    return r;
}

// Deposit the bits of k into the positions of `free` bits: the classic
// "insert a zero bit at position p" applied for each sorted fixed position.
__device__ __forceinline__ uint64_t insert_zero(uint64_t k, int p) {
    uint64_t lo = k & ((1ull << p) - 1ull);
    return ((k >> p) << (p + 1)) | lo;
}

// Up to 4 fixed bit positions, ascending, -1 terminated.
struct FixedBits {
    int pos[4];
    int count;
};

__device__ __forceinline__ uint64_t deposit_fixed(uint64_t k, const FixedBits& f) {
    for (int i = 0; i < f.count; ++i) k = insert_zero(k, f.pos[i]);
    return k;
}

// ---- device state -----------------------------------------------------------
constexpr int kMaxQubits = 34;

}  // namespace lq

struct lq_state {
    int n = 0;
    int dtype = LQ_C128;
    void* d_amps = nullptr;   // 2^n amplitudes
    uint64_t dim = 0;
    size_t amp_bytes = 16;
};

namespace lq {

// Thread-local error text for lq_last_error().
void set_error(const char* fmt, ...);
void clear_error();

// One library-wide lock taken at every C-ABI entry point.  All device work
// shares one stream, scratch buffers, plan buffers and the run-time compiled
// kernel cache, so calls from several threads (on distinct handles) run one
// at a time instead of racing.  Recursive: a host callback invoked by an
// optimiser entry point may call back into the library on the same thread.
std::recursive_mutex& api_mutex();
#define LQ_API_GUARD std::lock_guard<std::recursive_mutex> lq_api_guard_(::lq::api_mutex())

// The library's single HIP stream (default priority) and a small device
// scratch area for reductions, created on first use.
hipStream_t stream();
int ensure_init();

// Grows-on-demand device scratch (reduction partials, temporaries).  Calls
// are serialised by the library stream order; returns nullptr on failure.
void* scratch(size_t bytes);
// Second, independent scratch (e.g. a dense-matvec output vector).
void* scratch2(size_t bytes);

// Pinned host staging for small device-to-host results.
double* host_staging(size_t count);

#define LQ_HIP_CHECK(expr)                                                              \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess) {                                                         \
            ::lq::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e),      \
                            __FILE__, __LINE__);                                        \
            return LQ_ERR_DEVICE;                                                       \
        }                                                                               \
    } while (0)

#define LQ_LAUNCH_CHECK()                                                               \
    do {                                                                                \
        hipError_t _e = hipGetLastError();                                              \
        if (_e != hipSuccess) {                                                         \
            ::lq::set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(_e),  \
                            __FILE__, __LINE__);                                        \
            return LQ_ERR_DEVICE;                                                       \
        }                                                                               \
    } while (0)

// Validation helpers returning lq_status codes with the reference's messages.
int check_state(const lq_state* st);
int check_qubit(const lq_state* st, int q);

// ---- kernels exposed to other translation units ------------------------------
// Per-gate launchers (lq_gates.hip).
int launch_pair(lq_state* st, int target, uint64_t ctrl_mask, const double m[8], int op);
int launch_diag(lq_state* st, uint64_t fixed_mask, int sel0, int sel1, const double d[8]);
int launch_swap2(lq_state* st, int bit1, int bit2);
int launch_quad(lq_state* st, int bit_a, int bit_b, const double m[32]);
int launch_dense(lq_state* st, const double* d_matrix, uint64_t dim);

// Per-gate kernel variant (work units per wave, nontemporal access).
int set_gate_kernel_config(int unroll, int nontemporal);
int set_gate_unit_order(int classes);
int gate_unit_order();

// Reductions (lq_state.hip).
int reduce_norm_sqr(const lq_state* st, double* d_out);    // device scalar
int normalize_device(lq_state* st, double* norm_out);      // norm_out host, nullable

// Pair-kernel ops.
enum PairOp { kOpMatrix = 0, kOpSwap = 1, kOpNegY = 2 };

}  // namespace lq
