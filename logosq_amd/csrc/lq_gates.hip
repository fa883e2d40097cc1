// Per-gate kernels: the replacements of the reference's Gate::apply loops
// (src/gates/single_qubit.rs, two_qubit.rs, three_qubit.rs, parameterized.rs,
// src/gates/mod.rs:77-93).
//
// Layout: amplitudes contiguous, qubit q at bit (n-1-q).  A wavefront (64
// lanes) owns a "chunk" of 64 consecutive amplitudes, one per lane, so every
// global load and store is a coalesced 1 KiB (c128) / 512 B (c64) access.
//   - target bit >= 6: the partner amplitude lives in another chunk; the wave
//     loads both chunks and updates the pair in registers;
//   - target bit <  6: the partner lives in lane (lane ^ 2^t) of the same
//     chunk and is fetched with a cross-lane shuffle.
// Fixed bits (controls, the 1-half of phase gates) at or above bit 6 are
// deposited into the chunk index, so chunks that the gate does not touch are
// never read; fixed bits below 6 predicate the stores lane by lane.
#include <cmath>
#include <cstring>

#include "lq_common.h"

namespace lq {

constexpr int kWave = 64;
constexpr int kLaneBits = 6;
constexpr int kThreads = 256;

// Native two-element vectors for the nontemporal builtins.
template <typename T> struct nvec2;
template <> struct nvec2<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct nvec2<float> { typedef float type __attribute__((ext_vector_type(2))); };

// Global load/store of one amplitude; NT selects the nontemporal (streaming)
// form, which keeps a pass over a state far larger than L2 from evicting
// useful lines.
template <bool NT, typename V>
__device__ __forceinline__ V gload(const V* p) {
    using T = decltype(p->x);
    using N = typename nvec2<T>::type;
    N v;
    if constexpr (NT) v = __builtin_nontemporal_load(reinterpret_cast<const N*>(p));
    else v = *reinterpret_cast<const N*>(p);
    V r;
    r.x = v.x;
    r.y = v.y;
    return r;
}

template <bool NT, typename V>
__device__ __forceinline__ void gstore(V* p, V a) {
    using T = decltype(a.x);
    using N = typename nvec2<T>::type;
    N v;
    v.x = a.x;
    v.y = a.y;
    if constexpr (NT) __builtin_nontemporal_store(v, reinterpret_cast<N*>(p));
    else *reinterpret_cast<N*>(p) = v;
}

template <typename T> struct ChunkBits {
    int pos[4];  // fixed positions in chunk-index space (amp bit - 6), ascending
    int count;
};

__device__ __forceinline__ uint64_t deposit(uint64_t k, const int* pos, int count) {
    for (int i = 0; i < count; ++i) {
        int p = pos[i];
        k = ((k >> p) << (p + 1)) | (k & ((1ull << p) - 1ull));
    }
    return k;
}

template <typename V>
__device__ __forceinline__ V shfl_xor_c(V v, int mask) {
    V r;
    r.x = __shfl_xor(v.x, mask, kWave);
    r.y = __shfl_xor(v.y, mask, kWave);
    return r;
}

template <int OP, typename V, typename T>
__device__ __forceinline__ void pair_op(const Mat2<T>& m, V a0, V a1, V& o0, V& o1) {
    if (OP == kOpMatrix) {
        o0 = cadd(cmul(m.m[0], m.m[1], a0), cmul(m.m[2], m.m[3], a1));
        o1 = cadd(cmul(m.m[4], m.m[5], a0), cmul(m.m[6], m.m[7], a1));
    } else if (OP == kOpSwap) {
        o0 = a1;
        o1 = a0;
    } else if (OP == kOpNegY) {
        // YGate (single_qubit.rs:193-195): a0' = i*a1, a1' = -i*a0
        o0 = cmul((T)0, (T)1, a1);
        o1 = cmul((T)0, (T)-1, a0);
    } else {  // kOpH: s*(a0+a1), s*(a0-a1)  (single_qubit.rs:248-249)
        V s, d;
        s.x = a0.x + a1.x;
        s.y = a0.y + a1.y;
        d.x = a0.x - a1.x;
        d.y = a0.y - a1.y;
        o0 = cmul(m.m[0], (T)0, s);
        o1 = cmul(m.m[0], (T)0, d);
    }
}

constexpr int kOpHadamard = 3;

// Target bit >= 6: each work unit is a pair of chunks (target bit 0 / 1).
template <typename V, typename T, int OP, int U, bool NTM>
__global__ void __launch_bounds__(kThreads) k_pair_hi(V* __restrict__ psi, uint64_t units, ChunkBits<T> fb,
                                                      uint64_t chunk_or, uint64_t tstride, uint32_t lowctrl,
                                                      Mat2<T> m) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> kLaneBits;
    const uint64_t W = ((uint64_t)gridDim.x * blockDim.x) >> kLaneBits;
    const bool act = (lane & lowctrl) == lowctrl;
    V a0[U], a1[U];
    uint64_t idx[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
        uint64_t u = w + (uint64_t)j * W;
        idx[j] = ~0ull;
        if (u < units) {
            uint64_t c = deposit(u, fb.pos, fb.count) | chunk_or;
            idx[j] = (c << kLaneBits) | lane;
            a0[j] = gload<NTM>(psi + idx[j]);
            a1[j] = gload<NTM>(psi + idx[j] + tstride);
        }
    }
#pragma unroll
    for (int j = 0; j < U; ++j) {
        if (idx[j] != ~0ull && act) {
            V o0, o1;
            pair_op<OP, V, T>(m, a0[j], a1[j], o0, o1);
            gstore<NTM>(psi + idx[j], o0);
            gstore<NTM>(psi + idx[j] + tstride, o1);
        }
    }
}

// Target bit < 6: each work unit is one chunk; partner via lane shuffle.
template <typename V, typename T, int OP, int U, bool NTM>
__global__ void __launch_bounds__(kThreads) k_pair_lo(V* __restrict__ psi, uint64_t units, ChunkBits<T> fb,
                                                      uint64_t chunk_or, int tbit, uint32_t lowctrl,
                                                      uint32_t lane_limit, Mat2<T> m) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> kLaneBits;
    const uint64_t W = ((uint64_t)gridDim.x * blockDim.x) >> kLaneBits;
    const bool in_range = lane < lane_limit;
    const bool act = in_range && ((lane & lowctrl) == lowctrl);
    const bool hi_half = (lane >> tbit) & 1u;
    const int pmask = 1 << tbit;
    V a[U];
    uint64_t idx[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
        uint64_t u = w + (uint64_t)j * W;  // wave-uniform
        idx[j] = ~0ull;
        a[j].x = 0;
        a[j].y = 0;
        if (u < units) {
            uint64_t c = deposit(u, fb.pos, fb.count) | chunk_or;
            idx[j] = (c << kLaneBits) | lane;
            if (in_range) a[j] = gload<NTM>(psi + idx[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < U; ++j) {
        V p = shfl_xor_c(a[j], pmask);  // all lanes participate
        if (idx[j] != ~0ull && act) {
            V o0, o1;
            if (hi_half) {
                pair_op<OP, V, T>(m, p, a[j], o0, o1);
                gstore<NTM>(psi + idx[j], o1);
            } else {
                pair_op<OP, V, T>(m, a[j], p, o0, o1);
                gstore<NTM>(psi + idx[j], o0);
            }
        }
    }
}

// Diagonal phases: every amplitude whose fixed bits are all 1 is multiplied
// by d[bit(sel0) | bit(sel1) << 1].
template <typename V, typename T, int U, bool NTM>
__global__ void __launch_bounds__(kThreads) k_diag(V* __restrict__ psi, uint64_t units, ChunkBits<T> fb,
                                                   uint64_t chunk_or, uint32_t lowfixed, uint32_t lane_limit,
                                                   int sel0, int sel1, Diag4<T> d) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> kLaneBits;
    const uint64_t W = ((uint64_t)gridDim.x * blockDim.x) >> kLaneBits;
    const bool act = lane < lane_limit && ((lane & lowfixed) == lowfixed);
    V a[U];
    uint64_t idx[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
        uint64_t u = w + (uint64_t)j * W;
        idx[j] = ~0ull;
        if (u < units && act) {
            uint64_t c = deposit(u, fb.pos, fb.count) | chunk_or;
            idx[j] = (c << kLaneBits) | lane;
            a[j] = gload<NTM>(psi + idx[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < U; ++j) {
        if (idx[j] != ~0ull) {
            uint64_t i = idx[j];
            int k = 0;
            if (sel0 >= 0) k |= (int)((i >> sel0) & 1ull);
            if (sel1 >= 0) k |= (int)((i >> sel1) & 1ull) << 1;
            const T pr = k == 3 ? d.d[6] : k == 2 ? d.d[4] : k == 1 ? d.d[2] : d.d[0];
            const T pi = k == 3 ? d.d[7] : k == 2 ? d.d[5] : k == 1 ? d.d[3] : d.d[1];
            gstore<NTM>(psi + i, cmul(pr, pi, a[j]));
        }
    }
}

// SWAP (two_qubit.rs:40-62): swap |..1..0..> with |..0..1..> for bits (b1, b2).
template <typename V>
__global__ void __launch_bounds__(kThreads) k_swap2(V* __restrict__ psi, uint64_t count, int lo, int hi,
                                                    uint64_t m1, uint64_t m2) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += stride) {
        uint64_t i = insert_zero(insert_zero(k, lo), hi) | m1;
        uint64_t j = i ^ m1 ^ m2;
        V x = psi[i];
        V y = psi[j];
        psi[i] = y;
        psi[j] = x;
    }
}

// General 4x4 block on bits (ba, bb), sub-index (bit(ba) << 1) | bit(bb).
template <typename V, typename T>
__global__ void __launch_bounds__(kThreads) k_quad(V* __restrict__ psi, uint64_t count, int lo, int hi,
                                                   uint64_t ma, uint64_t mb, Mat4<T> m) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += stride) {
        uint64_t i = insert_zero(insert_zero(k, lo), hi);
        uint64_t ix[4] = {i, i | mb, i | ma, i | ma | mb};
        V a[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) a[s] = psi[ix[s]];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            V acc = cmul(m.m[8 * r], m.m[8 * r + 1], a[0]);
#pragma unroll
            for (int c = 1; c < 4; ++c) acc = cadd(acc, cmul(m.m[8 * r + 2 * c], m.m[8 * r + 2 * c + 1], a[c]));
            psi[ix[r]] = acc;
        }
    }
}

// MatrixGate::apply's dense matvec: out[i] = sum_j M[i][j] * psi[j]; one
// wavefront per row, computed in f64.
template <typename V>
__global__ void __launch_bounds__(kThreads) k_dense(const double2* __restrict__ mat, const V* __restrict__ psi,
                                                    V* __restrict__ out, uint64_t dim) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    uint64_t row = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> kLaneBits;
    if (row >= dim) return;
    const double2* mr = mat + row * dim;
    double re = 0.0, im = 0.0;
    for (uint64_t j = lane; j < dim; j += kWave) {
        double2 e = mr[j];
        V x = psi[j];
        re += e.x * (double)x.x - e.y * (double)x.y;
        im += e.x * (double)x.y + e.y * (double)x.x;
    }
    for (int off = kWave / 2; off > 0; off >>= 1) {
        re += __shfl_xor(re, off, kWave);
        im += __shfl_xor(im, off, kWave);
    }
    if (lane == 0) {
        V r;
        r.x = (decltype(r.x))re;
        r.y = (decltype(r.y))im;
        out[row] = r;
    }
}

// ---- host launchers -----------------------------------------------------------------
// Variant knobs (autotuned): work units per wave (2, 4, 8) and nontemporal
// global access.
static int g_unroll = 2;
static bool g_nontemporal = true;

struct Split {
    ChunkBits<double> fb{};   // chunk-space fixed positions (positions only)
    uint64_t chunk_or = 0;    // chunk-space value of the fixed bits (all 1)
    uint32_t lowmask = 0;     // lane-space fixed bits (must be 1)
    int chunk_bits = 0;       // number of chunk-index bits (n - 6, >= 0)
};

// Split a set of amplitude bits that must equal 1 (`ones`) and bits that are
// enumerated separately (`zeros`, e.g. a high target) into chunk/lane parts.
static Split split_bits(int n, uint64_t ones, uint64_t zeros) {
    Split s;
    s.chunk_bits = n > kLaneBits ? n - kLaneBits : 0;
    uint64_t all = ones | zeros;
    s.fb.count = 0;
    for (int b = kLaneBits; b < n; ++b)
        if (all & (1ull << b)) s.fb.pos[s.fb.count++] = b - kLaneBits;
    s.chunk_or = (ones >> kLaneBits);
    s.lowmask = (uint32_t)(ones & ((1ull << kLaneBits) - 1ull));
    return s;
}

static unsigned blocks_for_units(uint64_t units, int unroll) {
    uint64_t waves = (units + unroll - 1) / unroll;
    uint64_t blocks = (waves * kWave + kThreads - 1) / kThreads;
    if (blocks < 1) blocks = 1;
    return (unsigned)blocks;
}

template <typename V, typename T, int OP, int U, bool NTM>
static void pair_go(V* psi, bool hi, uint64_t units, const ChunkBits<T>& fb, uint64_t chunk_or, int tbit,
                    uint32_t lowmask, uint32_t lane_limit, const Mat2<T>& m) {
    unsigned g = blocks_for_units(units, U);
    if (hi)
        hipLaunchKernelGGL((k_pair_hi<V, T, OP, U, NTM>), dim3(g), dim3(kThreads), 0, stream(), psi, units, fb,
                           chunk_or, 1ull << tbit, lowmask, m);
    else
        hipLaunchKernelGGL((k_pair_lo<V, T, OP, U, NTM>), dim3(g), dim3(kThreads), 0, stream(), psi, units, fb,
                           chunk_or, tbit, lowmask, lane_limit, m);
}

template <typename V, typename T, int OP>
static void pair_pick(V* psi, bool hi, uint64_t units, const ChunkBits<T>& fb, uint64_t chunk_or, int tbit,
                      uint32_t lowmask, uint32_t lane_limit, const Mat2<T>& m) {
#define LQ_PAIR(U, NTM) pair_go<V, T, OP, U, NTM>(psi, hi, units, fb, chunk_or, tbit, lowmask, lane_limit, m)
    if (g_nontemporal) {
        if (g_unroll == 2) LQ_PAIR(2, true);
        else if (g_unroll == 8) LQ_PAIR(8, true);
        else LQ_PAIR(4, true);
    } else {
        if (g_unroll == 2) LQ_PAIR(2, false);
        else if (g_unroll == 8) LQ_PAIR(8, false);
        else LQ_PAIR(4, false);
    }
#undef LQ_PAIR
}

template <typename T, typename V>
static int launch_pair_t(lq_state* st, int tbit, uint64_t ctrl, const double md[8], int op) {
    Mat2<T> m;
    for (int i = 0; i < 8; ++i) m.m[i] = (T)md[i];
    const int n = st->n;
    V* psi = (V*)st->d_amps;
    const bool hi = tbit >= kLaneBits;
    Split s = split_bits(n, ctrl, hi ? (1ull << tbit) : 0);
    ChunkBits<T> fb;
    fb.count = s.fb.count;
    for (int i = 0; i < 4; ++i) fb.pos[i] = s.fb.pos[i];
    uint64_t units = (1ull << s.chunk_bits) >> s.fb.count;
    uint32_t lane_limit = n >= kLaneBits ? (uint32_t)kWave : (1u << n);
    switch (op) {
        case kOpMatrix: pair_pick<V, T, kOpMatrix>(psi, hi, units, fb, s.chunk_or, tbit, s.lowmask, lane_limit, m); break;
        case kOpSwap: pair_pick<V, T, kOpSwap>(psi, hi, units, fb, s.chunk_or, tbit, s.lowmask, lane_limit, m); break;
        case kOpNegY: pair_pick<V, T, kOpNegY>(psi, hi, units, fb, s.chunk_or, tbit, s.lowmask, lane_limit, m); break;
        default: pair_pick<V, T, kOpHadamard>(psi, hi, units, fb, s.chunk_or, tbit, s.lowmask, lane_limit, m); break;
    }
    LQ_LAUNCH_CHECK();
    return LQ_OK;
}

int launch_pair(lq_state* st, int tbit, uint64_t ctrl_mask, const double m[8], int op) {
    if (st->dtype == LQ_C128) return launch_pair_t<double, double2>(st, tbit, ctrl_mask, m, op);
    return launch_pair_t<float, float2>(st, tbit, ctrl_mask, m, op);
}

template <typename V, typename T, int U, bool NTM>
static void diag_go(V* psi, uint64_t units, const ChunkBits<T>& fb, uint64_t chunk_or, uint32_t lowmask,
                    uint32_t lane_limit, int sel0, int sel1, const Diag4<T>& d) {
    unsigned g = blocks_for_units(units, U);
    hipLaunchKernelGGL((k_diag<V, T, U, NTM>), dim3(g), dim3(kThreads), 0, stream(), psi, units, fb, chunk_or,
                       lowmask, lane_limit, sel0, sel1, d);
}

template <typename T, typename V>
static int launch_diag_t(lq_state* st, uint64_t fixed, int sel0, int sel1, const double dd[8]) {
    Diag4<T> d;
    for (int i = 0; i < 8; ++i) d.d[i] = (T)dd[i];
    Split s = split_bits(st->n, fixed, 0);
    ChunkBits<T> fb;
    fb.count = s.fb.count;
    for (int i = 0; i < 4; ++i) fb.pos[i] = s.fb.pos[i];
    uint64_t units = (1ull << s.chunk_bits) >> s.fb.count;
    uint32_t lane_limit = st->n >= kLaneBits ? (uint32_t)kWave : (1u << st->n);
    V* psi = (V*)st->d_amps;
#define LQ_DIAG(U, NTM) diag_go<V, T, U, NTM>(psi, units, fb, s.chunk_or, s.lowmask, lane_limit, sel0, sel1, d)
    if (g_nontemporal) {
        if (g_unroll == 2) LQ_DIAG(2, true);
        else if (g_unroll == 8) LQ_DIAG(8, true);
        else LQ_DIAG(4, true);
    } else {
        if (g_unroll == 2) LQ_DIAG(2, false);
        else if (g_unroll == 8) LQ_DIAG(8, false);
        else LQ_DIAG(4, false);
    }
#undef LQ_DIAG
    LQ_LAUNCH_CHECK();
    return LQ_OK;
}

int launch_diag(lq_state* st, uint64_t fixed_mask, int sel0, int sel1, const double d[8]) {
    if (st->dtype == LQ_C128) return launch_diag_t<double, double2>(st, fixed_mask, sel0, sel1, d);
    return launch_diag_t<float, float2>(st, fixed_mask, sel0, sel1, d);
}

int set_gate_kernel_config(int unroll, int nontemporal) {
    if (unroll != 2 && unroll != 4 && unroll != 8) {
        set_error("unroll must be 2, 4 or 8");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    g_unroll = unroll;
    g_nontemporal = nontemporal != 0;
    return LQ_OK;
}

static unsigned grid_stride_blocks(uint64_t count) {
    uint64_t b = (count + kThreads - 1) / kThreads;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return (unsigned)b;
}

int launch_swap2(lq_state* st, int b1, int b2) {
    int lo = b1 < b2 ? b1 : b2, hi = b1 < b2 ? b2 : b1;
    uint64_t count = st->dim >> 2;
    unsigned g = grid_stride_blocks(count);
    if (st->dtype == LQ_C128)
        hipLaunchKernelGGL(k_swap2<double2>, dim3(g), dim3(kThreads), 0, stream(), (double2*)st->d_amps, count, lo,
                           hi, 1ull << b1, 1ull << b2);
    else
        hipLaunchKernelGGL(k_swap2<float2>, dim3(g), dim3(kThreads), 0, stream(), (float2*)st->d_amps, count, lo, hi,
                           1ull << b1, 1ull << b2);
    LQ_LAUNCH_CHECK();
    return LQ_OK;
}

int launch_quad(lq_state* st, int ba, int bb, const double m[32]) {
    int lo = ba < bb ? ba : bb, hi = ba < bb ? bb : ba;
    uint64_t count = st->dim >> 2;
    unsigned g = grid_stride_blocks(count);
    if (st->dtype == LQ_C128) {
        Mat4<double> mm;
        for (int i = 0; i < 32; ++i) mm.m[i] = m[i];
        hipLaunchKernelGGL((k_quad<double2, double>), dim3(g), dim3(kThreads), 0, stream(), (double2*)st->d_amps,
                           count, lo, hi, 1ull << ba, 1ull << bb, mm);
    } else {
        Mat4<float> mm;
        for (int i = 0; i < 32; ++i) mm.m[i] = (float)m[i];
        hipLaunchKernelGGL((k_quad<float2, float>), dim3(g), dim3(kThreads), 0, stream(), (float2*)st->d_amps, count,
                           lo, hi, 1ull << ba, 1ull << bb, mm);
    }
    LQ_LAUNCH_CHECK();
    return LQ_OK;
}

int launch_dense(lq_state* st, const double* d_matrix, uint64_t dim) {
    void* out = scratch2(dim * st->amp_bytes);
    if (!out) {
        set_error("scratch allocation failed");
        return LQ_ERR_OUT_OF_MEMORY;
    }
    uint64_t blocks = (dim * kWave + kThreads - 1) / kThreads;
    if (st->dtype == LQ_C128)
        hipLaunchKernelGGL(k_dense<double2>, dim3((unsigned)blocks), dim3(kThreads), 0, stream(),
                           (const double2*)d_matrix, (const double2*)st->d_amps, (double2*)out, dim);
    else
        hipLaunchKernelGGL(k_dense<float2>, dim3((unsigned)blocks), dim3(kThreads), 0, stream(),
                           (const double2*)d_matrix, (const float2*)st->d_amps, (float2*)out, dim);
    LQ_LAUNCH_CHECK();
    LQ_HIP_CHECK(hipMemcpyAsync(st->d_amps, out, dim * st->amp_bytes, hipMemcpyDeviceToDevice, stream()));
    return normalize_device(st, nullptr);
}

}  // namespace lq
