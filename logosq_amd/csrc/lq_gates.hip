// Per-gate kernels: the replacements of the reference's Gate::apply loops
// (src/gates/single_qubit.rs, two_qubit.rs, three_qubit.rs, parameterized.rs,
// src/gates/mod.rs:77-93).
//
// Layout: amplitudes contiguous, qubit q at bit (n-1-q).  A gate's bits split
// into fixed bits (controls and the 1-half of phase gates: always 1 in the
// amplitudes it touches), enumerated bits (a lane holds every value of
// them) and free bits.  Lane l of a wavefront covers the 6 lowest free bits,
// the work unit the remaining ones, so each global load or store of a
// wavefront reads the densest run of amplitudes the gate touches (1 KiB of
// consecutive c128 amplitudes when no fixed bit lies below bit 6) and
// amplitudes the gate leaves alone are never read.
//   - pair target beyond the 6 lowest free bits: enumerated; the lane loads
//     both halves of its pair and updates them in registers;
//   - pair target among them: the partner lives in lane (lane ^ 2^tl) of the
//     same wavefront and is fetched with a cross-lane shuffle.
// SWAP and 4x4 blocks enumerate their bits above the six lowest; 4x4 bits
// below bit 6 stay lane bits and are exchanged by shuffles.
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

// Bit positions that a launch's lanes and work units skip, ascending: the
// fixed bits (value 1 in every amplitude the gate touches: controls, the
// 1-half of phase gates) except those with runs shorter than kDenseBytes,
// and the bits whose values a lane enumerates itself (a pair kernel's high
// target, a SWAP's bits, a 4x4 block's bits above bit 5).  The remaining
// "free" bits are numbered from the bottom: lane l covers the 6 lowest free
// bits and the work-unit index the rest, so a wavefront's 64 amplitudes are
// the densest run of amplitudes the gate must touch.
constexpr int kMaxSkip = 5;
struct SkipBits {
    // seg[i]: the bits of a compressed free-bit index that move up by i
    // places (i skipped positions lie below them)
    uint64_t seg[kMaxSkip + 1];
    uint64_t ones;   // the skipped fixed bits (all 1)
    uint64_t dense;  // fixed bits kept among the lane bits (see geometry())
    uint32_t order;  // 1: work units in transposed order (unit_order)
    uint32_t pad;
};

// Compressed free-bit index -> amplitude index with the skipped bits 0.  A
// bit-wise map, so the lane and unit parts can be spread separately and
// OR-ed; branch-free, and on the scalar unit for the wave-uniform unit part.
__device__ __forceinline__ uint64_t spread(uint64_t k, const SkipBits& s) {
    uint64_t r = k & s.seg[0];
#pragma unroll
    for (int i = 1; i <= kMaxSkip; ++i) r |= (k & s.seg[i]) << i;
    return r;
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

// Wave index and wave count of a launch (U units per wave, strided by W).
// Wave indices are wave-uniform: readfirstlane puts them (and the unit
// address arithmetic) on the scalar unit.  Every launch uses kThreads.
#define LQ_WAVE_IDS                                                                          \
    const uint32_t lane = threadIdx.x & (kWave - 1);                                         \
    const uint64_t w = (uint64_t)__builtin_amdgcn_readfirstlane(                             \
        blockIdx.x * (uint32_t)(kThreads / kWave) + (threadIdx.x >> kLaneBits));              \
    const uint64_t W = (uint64_t)gridDim.x * (kThreads / kWave);
#define LQ_UNIT(j) unit_order(w + (uint64_t)(j) * W, sk.order)

// Work unit -> the unit a wave processes.  order 1 swaps unit bits 0-4 with
// bits 5-9 (a 32 x 32 transpose inside each block of 1024 units; the host
// sets it only when the unit count is a multiple of 1024, so the map is a
// bijection on [0, units) and on the indices past it): 32 consecutive waves
// then take units 32 apart instead of neighbours.  At 30 qubits (complex128)
// this raises every kernel class: the in-place pair update with an
// enumerated target loses up to a fifth of its rate in the plain order at
// some partner distances (the stores of concurrent waves collide), and the
// transposed order runs it at 6.4 instead of 6.1 TB/s averaged over the
// targets; diagonal, SWAP and 4x4 kernels gain 2-8 %.
__device__ __forceinline__ uint64_t unit_order(uint64_t u, uint32_t order) {
    return order ? (u & ~1023ull) | ((u & 31ull) << 5) | ((u >> 5) & 31ull) : u;
}

// Pair kernel, target among the skipped bits: each lane holds both halves
// of its pair (a0 at idx, a1 at idx + tstride).
template <typename V, typename T, int OP, int U, bool NTM>
__global__ void __launch_bounds__(kThreads) k_pair_hi(V* __restrict__ psi, uint64_t units, SkipBits sk,
                                                      uint64_t tstride, uint32_t lane_limit, Mat2<T> m) {
    LQ_WAVE_IDS
    const uint64_t loff = spread(lane, sk) | sk.ones;
    const bool live = lane < lane_limit;
    const bool on = (loff & sk.dense) == sk.dense;
    V a0[U], a1[U];
    uint64_t idx[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
        const uint64_t u = LQ_UNIT(j);
        idx[j] = ~0ull;
        if (u < units && live) {
            idx[j] = spread(u << kLaneBits, sk) | loff;
            a0[j] = gload<NTM>(psi + idx[j]);
            a1[j] = gload<NTM>(psi + idx[j] + tstride);
        }
    }
#pragma unroll
    for (int j = 0; j < U; ++j) {
        if (idx[j] != ~0ull) {
            V o0 = a0[j], o1 = a1[j];
            if (on) pair_op<OP, V, T>(m, a0[j], a1[j], o0, o1);
            gstore<NTM>(psi + idx[j], o0);
            gstore<NTM>(psi + idx[j] + tstride, o1);
        }
    }
}

// Pair kernel, target among the 6 lowest free bits (lane bit tl): one
// amplitude per lane, partner through a cross-lane shuffle.
template <typename V, typename T, int OP, int U, bool NTM>
__global__ void __launch_bounds__(kThreads) k_pair_lo(V* __restrict__ psi, uint64_t units, SkipBits sk, int tl,
                                                      uint32_t lane_limit, Mat2<T> m) {
    LQ_WAVE_IDS
    const uint64_t loff = spread(lane, sk) | sk.ones;
    const bool live = lane < lane_limit;
    const bool on = (loff & sk.dense) == sk.dense;  // the partner lane agrees
    const bool hi_half = (lane >> tl) & 1u;
    const int pmask = 1 << tl;
    V a[U];
    uint64_t idx[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
        const uint64_t u = LQ_UNIT(j);  // wave-uniform
        idx[j] = ~0ull;
        a[j].x = 0;
        a[j].y = 0;
        if (u < units) {
            idx[j] = spread(u << kLaneBits, sk) | loff;
            if (live) a[j] = gload<NTM>(psi + idx[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < U; ++j) {
        V p = shfl_xor_c(a[j], pmask);  // every lane takes part
        if (idx[j] != ~0ull && live) {
            V o0, o1, r = a[j];
            if (hi_half) pair_op<OP, V, T>(m, p, a[j], o0, o1);
            else pair_op<OP, V, T>(m, a[j], p, o0, o1);
            // component-wise selects: a select of whole vector structs is
            // lowered through scratch memory
            if (on) {
                r.x = hi_half ? o1.x : o0.x;
                r.y = hi_half ? o1.y : o0.y;
            }
            gstore<NTM>(psi + idx[j], r);
        }
    }
}

// Diagonal phases: every amplitude whose fixed bits are all 1 is multiplied
// by d[bit(sel0) | bit(sel1) << 1].
template <typename V, typename T, int U, bool NTM>
__global__ void __launch_bounds__(kThreads) k_diag(V* __restrict__ psi, uint64_t units, SkipBits sk,
                                                   uint32_t lane_limit, int sel0, int sel1, Diag4<T> d) {
    LQ_WAVE_IDS
    const uint64_t loff = spread(lane, sk) | sk.ones;
    const bool live = lane < lane_limit;
    const bool on = (loff & sk.dense) == sk.dense;
    V a[U];
    uint64_t idx[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
        const uint64_t u = LQ_UNIT(j);
        idx[j] = ~0ull;
        if (u < units && live) {
            idx[j] = spread(u << kLaneBits, sk) | loff;
            a[j] = gload<NTM>(psi + idx[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < U; ++j) {
        if (idx[j] != ~0ull) {
            const uint64_t i = idx[j];
            int k = 0;
            if (sel0 >= 0) k |= (int)((i >> sel0) & 1ull);
            if (sel1 >= 0) k |= (int)((i >> sel1) & 1ull) << 1;
            const T pr = k == 3 ? d.d[6] : k == 2 ? d.d[4] : k == 1 ? d.d[2] : d.d[0];
            const T pi = k == 3 ? d.d[7] : k == 2 ? d.d[5] : k == 1 ? d.d[3] : d.d[1];
            V r = a[j];
            if (on) r = cmul(pr, pi, a[j]);
            gstore<NTM>(psi + i, r);
        }
    }
}

// SWAP (two_qubit.rs:40-62) of bits (ba, bb), both skipped: each lane
// exchanges |..ba=1..bb=0..> with |..ba=0..bb=1..>; the 00 and 11 amplitudes
// are never read.
template <typename V, int U, bool NTM>
__global__ void __launch_bounds__(kThreads) k_swap2(V* __restrict__ psi, uint64_t units, SkipBits sk,
                                                    uint32_t lane_limit, uint64_t ma, uint64_t mb) {
    LQ_WAVE_IDS
    const uint64_t loff = spread(lane, sk) | sk.ones;
    const bool live = lane < lane_limit;
    V x[U], y[U];
    uint64_t idx[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
        const uint64_t u = LQ_UNIT(j);
        idx[j] = ~0ull;
        if (u < units && live) {
            idx[j] = spread(u << kLaneBits, sk) | loff;
            x[j] = gload<NTM>(psi + (idx[j] | ma));
            y[j] = gload<NTM>(psi + (idx[j] | mb));
        }
    }
#pragma unroll
    for (int j = 0; j < U; ++j) {
        if (idx[j] != ~0ull) {
            gstore<NTM>(psi + (idx[j] | ma), y[j]);
            gstore<NTM>(psi + (idx[j] | mb), x[j]);
        }
    }
}

// General 4x4 block on bits (ba, bb), both skipped; sub-index
// (bit(ba) << 1) | bit(bb).  Each lane holds its four amplitudes.
template <typename V, typename T, int U, bool NTM>
__global__ void __launch_bounds__(kThreads) k_quad(V* __restrict__ psi, uint64_t units, SkipBits sk,
                                                   uint32_t lane_limit, uint64_t ma, uint64_t mb, Mat4<T> m) {
    LQ_WAVE_IDS
    const uint64_t loff = spread(lane, sk) | sk.ones;
    const bool live = lane < lane_limit;
    const uint64_t off[4] = {0, mb, ma, ma | mb};
    V a[U][4];
    uint64_t idx[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
        const uint64_t u = LQ_UNIT(j);
        idx[j] = ~0ull;
        if (u < units && live) {
            idx[j] = spread(u << kLaneBits, sk) | loff;
#pragma unroll
            for (int s = 0; s < 4; ++s) a[j][s] = gload<NTM>(psi + (idx[j] | off[s]));
        }
    }
#pragma unroll
    for (int j = 0; j < U; ++j) {
        if (idx[j] != ~0ull) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                V acc = cmul(m.m[8 * r], m.m[8 * r + 1], a[j][0]);
#pragma unroll
                for (int c = 1; c < 4; ++c) acc = cadd(acc, cmul(m.m[8 * r + 2 * c], m.m[8 * r + 2 * c + 1], a[j][c]));
                gstore<NTM>(psi + (idx[j] | off[r]), acc);
            }
        }
    }
}

// 4x4 block with one (NLO = 1) or both (NLO = 2) of its bits among the six
// lowest bits.  Those are lane bits (la for bit a, lb for bit b; -1 for an
// enumerated bit, whose mask is mh), so every load and store of a wavefront
// is one coalesced run; a lane's partners come through cross-lane shuffles
// and the lane computes its own outputs with its row of the block (selected
// per lane by its own bit values).  Sub-index (bit a << 1) | bit b.
template <typename T>
__device__ __forceinline__ T pick2(bool c, T x, T y) { return c ? y : x; }

template <typename V, typename T, int NLO, int U, bool NTM>
__global__ void __launch_bounds__(kThreads) k_quad_lo(V* __restrict__ psi, uint64_t units, SkipBits sk,
                                                      uint32_t lane_limit, int la, int lb, uint64_t mh,
                                                      Mat4<T> m) {
    LQ_WAVE_IDS
    const uint64_t loff = spread(lane, sk) | sk.ones;
    const bool live = lane < lane_limit;
    constexpr int NV = NLO == 2 ? 1 : 2;  // amplitudes per lane
    V v[U][NV];
    uint64_t idx[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
        const uint64_t u = LQ_UNIT(j);
        idx[j] = ~0ull;
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            v[j][k].x = 0;
            v[j][k].y = 0;
        }
        if (u < units) {
            idx[j] = spread(u << kLaneBits, sk) | loff;
            if (live) {
                v[j][0] = gload<NTM>(psi + idx[j]);
                if constexpr (NV == 2) v[j][1] = gload<NTM>(psi + (idx[j] | mh));
            }
        }
    }
    if constexpr (NLO == 2) {
        const int xa = (lane >> la) & 1, xb = (lane >> lb) & 1;
        const int so = (xa << 1) | xb;  // own sub-index
        // w_e = amplitude with sub-index so ^ e; its coefficient M[so][so ^ e]
        T cr[4], ci[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            T r4[4], i4[4];
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
                r4[s4] = m.m[8 * s4 + 2 * (s4 ^ e)];
                i4[s4] = m.m[8 * s4 + 2 * (s4 ^ e) + 1];
            }
            cr[e] = pick2(xa, pick2(xb, r4[0], r4[1]), pick2(xb, r4[2], r4[3]));
            ci[e] = pick2(xa, pick2(xb, i4[0], i4[1]), pick2(xb, i4[2], i4[3]));
        }
        const int mask[4] = {0, 1 << lb, 1 << la, (1 << la) | (1 << lb)};
#pragma unroll
        for (int j = 0; j < U; ++j) {
            V w[4];
            w[0] = v[j][0];
#pragma unroll
            for (int e = 1; e < 4; ++e) w[e] = shfl_xor_c(v[j][0], mask[e]);  // every lane takes part
            if (idx[j] != ~0ull && live) {
                V acc = cmul(cr[0], ci[0], w[0]);
#pragma unroll
                for (int e = 1; e < 4; ++e) acc = cadd(acc, cmul(cr[e], ci[e], w[e]));
                gstore<NTM>(psi + idx[j], acc);
            }
        }
    } else {
        const bool lo_is_a = la >= 0;
        const int lbit = lo_is_a ? la : lb;
        const int x = (lane >> lbit) & 1;  // own value of the lane bit
        // sub-index of (lane-bit value xl, enumerated-bit value yh)
        auto sub = [&](int xl, int yh) { return lo_is_a ? ((xl << 1) | yh) : ((yh << 1) | xl); };
        // coefficient of (same / other lane value, yh) in output row (x, yo)
        T cr[2][2][2], ci[2][2][2];
#pragma unroll
        for (int yo = 0; yo < 2; ++yo)
#pragma unroll
            for (int o = 0; o < 2; ++o)
#pragma unroll
                for (int yh = 0; yh < 2; ++yh) {
                    const int r0 = sub(0, yo), c0 = sub(o, yh);          // own lane value 0
                    const int r1 = sub(1, yo), c1 = sub(1 - o, yh);      // own lane value 1
                    cr[yo][o][yh] = pick2(x, m.m[8 * r0 + 2 * c0], m.m[8 * r1 + 2 * c1]);
                    ci[yo][o][yh] = pick2(x, m.m[8 * r0 + 2 * c0 + 1], m.m[8 * r1 + 2 * c1 + 1]);
                }
#pragma unroll
        for (int j = 0; j < U; ++j) {
            const V p0 = shfl_xor_c(v[j][0], 1 << lbit), p1 = shfl_xor_c(v[j][1], 1 << lbit);
            if (idx[j] != ~0ull && live) {
#pragma unroll
                for (int yo = 0; yo < 2; ++yo) {
                    V acc = cmul(cr[yo][0][0], ci[yo][0][0], v[j][0]);
                    acc = cadd(acc, cmul(cr[yo][0][1], ci[yo][0][1], v[j][1]));
                    acc = cadd(acc, cmul(cr[yo][1][0], ci[yo][1][0], p0));
                    acc = cadd(acc, cmul(cr[yo][1][1], ci[yo][1][1], p1));
                    gstore<NTM>(psi + (yo ? (idx[j] | mh) : idx[j]), acc);
                }
            }
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
static int g_unroll = 1;
static bool g_nontemporal = true;
// Kernel classes whose units run in transposed order (LQ_UNITS_* bits).
constexpr int kUnitOrderDefault = LQ_UNITS_ALL;
static int g_unit_order = kUnitOrderDefault;

// Fixed bits whose runs of equal value are shorter than this many bytes
// stay among the lane bits (a lane whose copy of them is not all 1 stores
// its amplitude back unchanged): HBM writes that cover only part of such a
// short run cost more than writing the whole run.
#ifndef LQ_DENSE_BYTES
#define LQ_DENSE_BYTES 128
#endif
constexpr uint64_t kDenseBytes = LQ_DENSE_BYTES;

// True when bit b of an amplitude index splits the state into runs shorter
// than kDenseBytes.
static inline bool short_run(int b, size_t amp_bytes) { return ((uint64_t)amp_bytes << b) < kDenseBytes; }

// The launch geometry of one gate: skipped bits, the number of free bits
// (n - skipped), the work units (each 64 free-bit values) and the live
// lanes (fewer than 64 only for tiny states).
struct Geometry {
    SkipBits sk{};
    bool ok = true;  // at most kMaxSkip skipped bits
    int free_bits = 0;
    uint64_t units = 1;
    uint32_t lane_limit = kWave;
};

static uint64_t dense_bits(uint64_t ones, size_t amp_bytes) {
    uint64_t d = 0;
    for (int b = 0; b < 64 && short_run(b, amp_bytes); ++b) d |= ones & (1ull << b);
    return d;
}

static Geometry geometry(int n, uint64_t ones, uint64_t skip, size_t amp_bytes) {
    Geometry g;
    const uint64_t dense = dense_bits(ones, amp_bytes);
    const uint64_t all = (ones & ~dense) | skip;
    for (int i = 0; i <= kMaxSkip; ++i) g.sk.seg[i] = 0;
    int below = 0, j = 0;  // skipped positions passed; compressed bit index
    for (int b = 0; b < n; ++b) {
        if (all & (1ull << b)) {
            if (++below > kMaxSkip) {
                g.ok = false;
                return g;
            }
        } else {
            g.sk.seg[below] |= 1ull << j++;
        }
    }
    g.sk.ones = ones & ~dense;
    g.sk.dense = dense;
    g.sk.order = 0;
    g.sk.pad = 0;
    g.free_bits = j;
    if (g.free_bits >= kLaneBits) {
        g.units = 1ull << (g.free_bits - kLaneBits);
        g.lane_limit = kWave;
    } else {
        g.units = 1;
        g.lane_limit = 1u << g.free_bits;
    }
    return g;
}

// The geometry with the unit order of kernel class `cls` (LQ_UNITS_*).
static Geometry ordered(Geometry g, int cls) {
    g.sk.order = (g_unit_order & cls) && g.units >= 1024 && g.units % 1024 == 0 ? 1u : 0u;
    return g;
}

static unsigned blocks_for_units(uint64_t units, int unroll) {
    uint64_t waves = (units + unroll - 1) / unroll;
    uint64_t blocks = (waves * kWave + kThreads - 1) / kThreads;
    if (blocks < 1) blocks = 1;
    return (unsigned)blocks;
}

// Calls F.template operator()<U, NT>() for the configured variant.  The
// SWAP and 4x4 kernels ask for at least 2 units per wavefront (min_unroll):
// with their two or four loads per lane they are faster with two units
// than with one, which suits the pair and diagonal kernels.
template <typename F>
static void with_variant(F&& f, int min_unroll = 1) {
    const int u = g_unroll < min_unroll ? min_unroll : g_unroll;
    if (g_nontemporal) {
        if (u == 1) f.template operator()<1, true>();
        else if (u == 2) f.template operator()<2, true>();
        else if (u == 8) f.template operator()<8, true>();
        else f.template operator()<4, true>();
    } else {
        if (u == 1) f.template operator()<1, false>();
        else if (u == 2) f.template operator()<2, false>();
        else if (u == 8) f.template operator()<8, false>();
        else f.template operator()<4, false>();
    }
}

template <typename V, typename T, int OP>
static void pair_launch(V* psi, const Geometry& g, bool hi, int tsel, const Mat2<T>& m) {
    with_variant([&]<int U, bool NT>() {
        const unsigned blocks = blocks_for_units(g.units, U);
        if (hi)
            hipLaunchKernelGGL((k_pair_hi<V, T, OP, U, NT>), dim3(blocks), dim3(kThreads), 0, stream(), psi, g.units,
                               g.sk, 1ull << tsel, g.lane_limit, m);
        else
            hipLaunchKernelGGL((k_pair_lo<V, T, OP, U, NT>), dim3(blocks), dim3(kThreads), 0, stream(), psi, g.units,
                               g.sk, tsel, g.lane_limit, m);
    });
}

template <typename T, typename V>
static int launch_pair_t(lq_state* st, int tbit, uint64_t ctrl, const double md[8], int op) {
    Mat2<T> m;
    for (int i = 0; i < 8; ++i) m.m[i] = (T)md[i];
    const int n = st->n;
    V* psi = (V*)st->d_amps;
    // The target's rank among the free bits (skipped controls left out):
    // below 6 the partner is in the same wavefront, reached by a shuffle on
    // lane bit tl.
    const uint64_t skipped = ctrl & ~dense_bits(ctrl, st->amp_bytes);
    const int tl = tbit - __builtin_popcountll(skipped & ((1ull << tbit) - 1ull));
    const bool hi = tl >= kLaneBits;
    const Geometry g = ordered(geometry(n, ctrl, hi ? (1ull << tbit) : 0, st->amp_bytes),
                               hi ? LQ_UNITS_PAIR_HI : LQ_UNITS_PAIR_LO);
    const int tsel = hi ? tbit : tl;
    if (!g.ok) {
        set_error("gate fixes more than %d bits", kMaxSkip);
        return LQ_ERR_INVALID_ARGUMENT;
    }
    switch (op) {
        case kOpMatrix: pair_launch<V, T, kOpMatrix>(psi, g, hi, tsel, m); break;
        case kOpSwap: pair_launch<V, T, kOpSwap>(psi, g, hi, tsel, m); break;
        case kOpNegY: pair_launch<V, T, kOpNegY>(psi, g, hi, tsel, m); break;
        default: pair_launch<V, T, kOpHadamard>(psi, g, hi, tsel, m); break;
    }
    LQ_LAUNCH_CHECK();
    return LQ_OK;
}

int launch_pair(lq_state* st, int tbit, uint64_t ctrl_mask, const double m[8], int op) {
    if (st->dtype == LQ_C128) return launch_pair_t<double, double2>(st, tbit, ctrl_mask, m, op);
    return launch_pair_t<float, float2>(st, tbit, ctrl_mask, m, op);
}

template <typename T, typename V>
static int launch_diag_t(lq_state* st, uint64_t fixed, int sel0, int sel1, const double dd[8]) {
    Diag4<T> d;
    for (int i = 0; i < 8; ++i) d.d[i] = (T)dd[i];
    const Geometry g = ordered(geometry(st->n, fixed, 0, st->amp_bytes), LQ_UNITS_DIAG);
    V* psi = (V*)st->d_amps;
    if (!g.ok) {
        set_error("gate fixes more than %d bits", kMaxSkip);
        return LQ_ERR_INVALID_ARGUMENT;
    }
    with_variant([&]<int U, bool NT>() {
        hipLaunchKernelGGL((k_diag<V, T, U, NT>), dim3(blocks_for_units(g.units, U)), dim3(kThreads), 0, stream(),
                           psi, g.units, g.sk, g.lane_limit, sel0, sel1, d);
    });
    LQ_LAUNCH_CHECK();
    return LQ_OK;
}

int launch_diag(lq_state* st, uint64_t fixed_mask, int sel0, int sel1, const double d[8]) {
    if (st->dtype == LQ_C128) return launch_diag_t<double, double2>(st, fixed_mask, sel0, sel1, d);
    return launch_diag_t<float, float2>(st, fixed_mask, sel0, sel1, d);
}

int set_gate_unit_order(int classes) {
    if (classes == -1) classes = kUnitOrderDefault;
    if (classes < 0 || classes > LQ_UNITS_ALL) {
        set_error("unit order classes must be a mask of LQ_UNITS_* bits (0..%d)", LQ_UNITS_ALL);
        return LQ_ERR_INVALID_ARGUMENT;
    }
    g_unit_order = classes;
    return LQ_OK;
}

int gate_unit_order() { return g_unit_order; }

int set_gate_kernel_config(int unroll, int nontemporal) {
    if (unroll != 1 && unroll != 2 && unroll != 4 && unroll != 8) {
        set_error("unroll must be 1, 2, 4 or 8");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    g_unroll = unroll;
    g_nontemporal = nontemporal != 0;
    return LQ_OK;
}

template <typename V>
static void swap2_t(lq_state* st, int ba, int bb) {
    const Geometry g = ordered(geometry(st->n, 0, (1ull << ba) | (1ull << bb), st->amp_bytes), LQ_UNITS_SWAP);
    V* psi = (V*)st->d_amps;
    with_variant([&]<int U, bool NT>() {
        hipLaunchKernelGGL((k_swap2<V, U, NT>), dim3(blocks_for_units(g.units, U)), dim3(kThreads), 0, stream(),
                           psi, g.units, g.sk, g.lane_limit, 1ull << ba, 1ull << bb);
    }, 2);
}

int launch_quad(lq_state* st, int ba, int bb, const double m[32]);

int launch_swap2(lq_state* st, int b1, int b2) {
    if (short_run(b1, st->amp_bytes) || short_run(b2, st->amp_bytes)) {
        // The 01 and 10 amplitudes form short runs: write every amplitude
        // (the 4x4 permutation, exact: products with 1 and 0 only).
        double pm[32] = {0};
        pm[0] = pm[8 * 1 + 2 * 2] = pm[8 * 2 + 2 * 1] = pm[8 * 3 + 2 * 3] = 1.0;
        return launch_quad(st, b1, b2, pm);
    }
    if (st->dtype == LQ_C128) swap2_t<double2>(st, b1, b2);
    else swap2_t<float2>(st, b1, b2);
    LQ_LAUNCH_CHECK();
    return LQ_OK;
}

template <typename V, typename T>
static void quad_t(lq_state* st, int ba, int bb, const double* md) {
    Mat4<T> m;
    for (int i = 0; i < 32; ++i) m.m[i] = (T)md[i];
    V* psi = (V*)st->d_amps;
    const bool lo_a = ba < kLaneBits, lo_b = bb < kLaneBits;
    if (!lo_a && !lo_b) {
        // both bits enumerated: four coalesced loads per lane
        const Geometry g = ordered(geometry(st->n, 0, (1ull << ba) | (1ull << bb), st->amp_bytes), LQ_UNITS_QUAD);
        with_variant([&]<int U, bool NT>() {
            hipLaunchKernelGGL((k_quad<V, T, U, NT>), dim3(blocks_for_units(g.units, U)), dim3(kThreads), 0,
                               stream(), psi, g.units, g.sk, g.lane_limit, 1ull << ba, 1ull << bb, m);
        }, 2);
    } else if (lo_a && lo_b) {
        // both among the lane bits (nothing skipped, so lane bit = amplitude bit)
        const Geometry g = ordered(geometry(st->n, 0, 0, st->amp_bytes), LQ_UNITS_QUAD_LO);
        with_variant([&]<int U, bool NT>() {
            hipLaunchKernelGGL((k_quad_lo<V, T, 2, U, NT>), dim3(blocks_for_units(g.units, U)), dim3(kThreads), 0,
                               stream(), psi, g.units, g.sk, g.lane_limit, ba, bb, (uint64_t)0, m);
        }, 2);
    } else {
        // one lane bit (below bit 6, so below the skipped one), one enumerated
        const int hb = lo_a ? bb : ba;
        const Geometry g = ordered(geometry(st->n, 0, 1ull << hb, st->amp_bytes), LQ_UNITS_QUAD_LO);
        with_variant([&]<int U, bool NT>() {
            hipLaunchKernelGGL((k_quad_lo<V, T, 1, U, NT>), dim3(blocks_for_units(g.units, U)), dim3(kThreads), 0,
                               stream(), psi, g.units, g.sk, g.lane_limit, lo_a ? ba : -1, lo_b ? bb : -1,
                               1ull << hb, m);
        }, 2);
    }
}

int launch_quad(lq_state* st, int ba, int bb, const double m[32]) {
    if (st->dtype == LQ_C128) quad_t<double2, double>(st, ba, bb, m);
    else quad_t<float2, float>(st, ba, bb, m);
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
