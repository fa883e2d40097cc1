// Device code of the grouped tile kernels: a thread holds a 16-amplitude
// subcube (all values of 4 tile bits) in registers and applies a group's ops
// to it.  Shared by the generic kernel (k_tileg, which dispatches on each
// op's kind and position at run time) and the pass kernels specialised at
// run time (which call the same templates with every position, kind and
// control mask fixed at compile time).  Self-contained: hipRTC compiles it
// from text.
#pragma once

#include "lq_tile_types.h"

namespace lq {

__device__ __forceinline__ uint32_t ins0(uint32_t k, int p) {
    return ((k >> p) << (p + 1)) | (k & ((1u << p) - 1u));
}

// LDS slot of tile element e: XOR-swizzles the low 4 bits with bits 4..7 so
// that 16-amplitude subcubes over any 4 bits, read by consecutive lanes, fall
// in different LDS banks.
__device__ __forceinline__ uint32_t swz(uint32_t e) { return e ^ ((e >> 4) & 15u); }

// One pair op on subcube position P.  `cin` lists the controls inside the
// subcube: for a compile-time slot s the test is uniform (a scalar branch),
// never a per-element select.
template <int P, int KIND, typename T>
__device__ __forceinline__ void g_pair(T* xr, T* xi, const TileOp& op, int cin) {
    const T m0 = (T)op.m[0], m1 = (T)op.m[1], m2 = (T)op.m[2], m3 = (T)op.m[3];
    const T m4 = (T)op.m[4], m5 = (T)op.m[5], m6 = (T)op.m[6], m7 = (T)op.m[7];
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        if (s & (1 << P)) continue;
        if ((s & cin) != cin) continue;
        const int t = s | (1 << P);
        const T ar = xr[s], ai = xi[s], br = xr[t], bi = xi[t];
        if (KIND == kTPairReal) {
            xr[s] = m0 * ar + m2 * br;
            xi[s] = m0 * ai + m2 * bi;
            xr[t] = m4 * ar + m6 * br;
            xi[t] = m4 * ai + m6 * bi;
        } else if (KIND == kTPairRX) {
            // [[c, i*m3], [i*m5, c']]: re/im of c*a + i*m3*b etc.
            xr[s] = m0 * ar - m3 * bi;
            xi[s] = m0 * ai + m3 * br;
            xr[t] = m6 * br - m5 * ai;
            xi[t] = m6 * bi + m5 * ar;
        } else if (KIND == kTPairMat) {
            xr[s] = (m0 * ar - m1 * ai) + (m2 * br - m3 * bi);
            xi[s] = (m0 * ai + m1 * ar) + (m2 * bi + m3 * br);
            xr[t] = (m4 * ar - m5 * ai) + (m6 * br - m7 * bi);
            xi[t] = (m4 * ai + m5 * ar) + (m6 * bi + m7 * br);
        } else if (KIND == kTPairSwap) {
            xr[s] = br;
            xi[s] = bi;
            xr[t] = ar;
            xi[t] = ai;
        } else if (KIND == kTPairNegY) {
            // a0' = i*a1, a1' = -i*a0 (single_qubit.rs:193-195)
            xr[s] = -bi;
            xi[s] = br;
            xr[t] = ai;
            xi[t] = -ar;
        } else {
            // s*(a0 + a1), s*(a0 - a1) (single_qubit.rs:248-249)
            xr[s] = m0 * (ar + br);
            xi[s] = m0 * (ai + bi);
            xr[t] = m0 * (ar - br);
            xi[t] = m0 * (ai - bi);
        }
    }
}

template <int KIND, typename T>
__device__ __forceinline__ void g_pair_at(T* xr, T* xi, const TileOp& op, int cin) {
    switch (op.sp0) {
        case 0: g_pair<0, KIND, T>(xr, xi, op, cin); break;
        case 1: g_pair<1, KIND, T>(xr, xi, op, cin); break;
        case 2: g_pair<2, KIND, T>(xr, xi, op, cin); break;
        default: g_pair<3, KIND, T>(xr, xi, op, cin); break;
    }
}

// SWAP of subcube positions P0 (value 1) and P1 (value 0).
template <int P0, int P1, typename T>
__device__ __forceinline__ void g_swap2(T* xr, T* xi) {
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        if (!((s >> P0) & 1) || ((s >> P1) & 1)) continue;
        const int t = s ^ (1 << P0) ^ (1 << P1);
        T r = xr[s], i = xi[s];
        xr[s] = xr[t];
        xi[s] = xi[t];
        xr[t] = r;
        xi[t] = i;
    }
}

// 4x4 block on subcube positions PA (bit a) and PB (bit b), sub-index
// (bit a << 1) | bit b.
template <int PA, int PB, typename T>
__device__ __forceinline__ void g_quad(T* xr, T* xi, const double* m) {
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        if (((s >> PA) & 1) || ((s >> PB) & 1)) continue;
        const int ix[4] = {s, s | (1 << PB), s | (1 << PA), s | (1 << PA) | (1 << PB)};
        T ar[4], ai[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            ar[c] = xr[ix[c]];
            ai[c] = xi[ix[c]];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            T yr = (T)0, yi = (T)0;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const T er = (T)m[8 * r + 2 * c], ei = (T)m[8 * r + 2 * c + 1];
                yr += er * ar[c] - ei * ai[c];
                yi += er * ai[c] + ei * ar[c];
            }
            xr[ix[r]] = yr;
            xi[ix[r]] = yi;
        }
    }
}

template <bool RICH, typename T>
__device__ __forceinline__ void g_apply(T* xr, T* xi, const TileOp& op, uint64_t gb) {
    // controls outside the subcube are fixed for the whole subcube
    if ((gb & op.cout) != op.cout) return;
    const int kind = op.kind;
    const int cin = op.cin;
    if (kind == kTDiag) {
        const int s0p = op.s0p, s1p = op.s1p;
        const int tb0 = op.osel0 >= 0 ? (int)((gb >> op.osel0) & 1ull) : 0;
        const int tb1 = op.osel1 >= 0 ? (int)((gb >> op.osel1) & 1ull) : 0;
        const T r0 = (T)op.m[0], i0 = (T)op.m[1], r1 = (T)op.m[2], i1 = (T)op.m[3];
        const T r2 = (T)op.m[4], i2 = (T)op.m[5], r3 = (T)op.m[6], i3 = (T)op.m[7];
        if (s0p < 0 && s1p < 0) {
            // selectors outside the subcube: one phase for the thread
            const T pr = tb1 ? (tb0 ? r3 : r2) : (tb0 ? r1 : r0);
            const T pi = tb1 ? (tb0 ? i3 : i2) : (tb0 ? i1 : i0);
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                if ((s & cin) != cin) continue;
                const T yr = pr * xr[s] - pi * xi[s];
                const T yi = pr * xi[s] + pi * xr[s];
                xr[s] = yr;
                xi[s] = yi;
            }
        } else if (s1p < 0) {
            // first selector inside: two candidate phases per thread
            const T p0r = tb1 ? r2 : r0, p0i = tb1 ? i2 : i0;
            const T p1r = tb1 ? r3 : r1, p1i = tb1 ? i3 : i1;
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                if ((s & cin) != cin) continue;
                const bool b = (s >> s0p) & 1;
                const T pr = b ? p1r : p0r, pi = b ? p1i : p0i;
                const T yr = pr * xr[s] - pi * xi[s];
                const T yi = pr * xi[s] + pi * xr[s];
                xr[s] = yr;
                xi[s] = yi;
            }
        } else {
            // general: second selector inside (first inside or out)
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                if ((s & cin) != cin) continue;
                const int b0 = s0p >= 0 ? ((s >> s0p) & 1) : tb0;
                const int b1 = (s >> s1p) & 1;
                const T pr = b1 ? (b0 ? r3 : r2) : (b0 ? r1 : r0);
                const T pi = b1 ? (b0 ? i3 : i2) : (b0 ? i1 : i0);
                const T yr = pr * xr[s] - pi * xi[s];
                const T yi = pr * xi[s] + pi * xr[s];
                xr[s] = yr;
                xi[s] = yi;
            }
        }
    } else if (kind == kTPairReal) {
        g_pair_at<kTPairReal, T>(xr, xi, op, cin);
    } else if (kind == kTPairRX) {
        g_pair_at<kTPairRX, T>(xr, xi, op, cin);
    } else if (kind == kTPairMat) {
        g_pair_at<kTPairMat, T>(xr, xi, op, cin);
    } else if (kind == kTPairSwap) {
        g_pair_at<kTPairSwap, T>(xr, xi, op, cin);
    } else if (kind == kTPairNegY) {
        g_pair_at<kTPairNegY, T>(xr, xi, op, cin);
    } else if (kind == kTPairH) {
        g_pair_at<kTPairH, T>(xr, xi, op, cin);
    } else if constexpr (RICH) {
        const int code = op.sp0 * 4 + op.sp1;
        if (kind == kTSwap2) {
            switch (code) {
#define LQ_SW(a, b) case a * 4 + b: g_swap2<a, b, T>(xr, xi); break;
                LQ_SW(0, 1) LQ_SW(0, 2) LQ_SW(0, 3) LQ_SW(1, 0) LQ_SW(1, 2) LQ_SW(1, 3)
                LQ_SW(2, 0) LQ_SW(2, 1) LQ_SW(2, 3) LQ_SW(3, 0) LQ_SW(3, 1) LQ_SW(3, 2)
#undef LQ_SW
                default: break;
            }
        } else {
            switch (code) {
#define LQ_QD(a, b) case a * 4 + b: g_quad<a, b, T>(xr, xi, op.m); break;
                LQ_QD(0, 1) LQ_QD(0, 2) LQ_QD(0, 3) LQ_QD(1, 0) LQ_QD(1, 2) LQ_QD(1, 3)
                LQ_QD(2, 0) LQ_QD(2, 1) LQ_QD(2, 3) LQ_QD(3, 0) LQ_QD(3, 1) LQ_QD(3, 2)
#undef LQ_QD
                default: break;
            }
        }
    }
}

// ---- kernels specialised at run time -------------------------------------------
// The same updates with the op's subcube position, kind and in-subcube
// control mask CINV known at compile time: every slot's test folds away, so
// an op is straight-line arithmetic on fixed registers.
template <int P, int KIND, int CINV, typename T>
__device__ __forceinline__ void j_pair(T* xr, T* xi, const TileOp& op) {
    const T m0 = (T)op.m[0], m1 = (T)op.m[1], m2 = (T)op.m[2], m3 = (T)op.m[3];
    const T m4 = (T)op.m[4], m5 = (T)op.m[5], m6 = (T)op.m[6], m7 = (T)op.m[7];
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        if (s & (1 << P)) continue;
        if ((s & CINV) != CINV) continue;
        const int t = s | (1 << P);
        const T ar = xr[s], ai = xi[s], br = xr[t], bi = xi[t];
        if (KIND == kTPairReal) {
            xr[s] = m0 * ar + m2 * br;
            xi[s] = m0 * ai + m2 * bi;
            xr[t] = m4 * ar + m6 * br;
            xi[t] = m4 * ai + m6 * bi;
        } else if (KIND == kTPairRX) {
            xr[s] = m0 * ar - m3 * bi;
            xi[s] = m0 * ai + m3 * br;
            xr[t] = m6 * br - m5 * ai;
            xi[t] = m6 * bi + m5 * ar;
        } else if (KIND == kTPairMat) {
            xr[s] = (m0 * ar - m1 * ai) + (m2 * br - m3 * bi);
            xi[s] = (m0 * ai + m1 * ar) + (m2 * bi + m3 * br);
            xr[t] = (m4 * ar - m5 * ai) + (m6 * br - m7 * bi);
            xi[t] = (m4 * ai + m5 * ar) + (m6 * bi + m7 * br);
        } else if (KIND == kTPairSwap) {
            xr[s] = br;
            xi[s] = bi;
            xr[t] = ar;
            xi[t] = ai;
        } else if (KIND == kTPairNegY) {
            xr[s] = -bi;
            xi[s] = br;
            xr[t] = ai;
            xi[t] = -ar;
        } else {
            xr[s] = m0 * (ar + br);
            xi[s] = m0 * (ai + bi);
            xr[t] = m0 * (ar - br);
            xi[t] = m0 * (ai - bi);
        }
    }
}

// Diagonal op: phase m[2k], m[2k+1] with k = b0 | b1 << 1, where b0 (b1) is
// the slot's bit at subcube position S0 (S1) or, for -1, the thread's
// outside value tb0 (tb1); slots whose in-subcube controls CINV are not all
// 1 keep their value.
template <int S0, int S1, int CINV, typename T>
__device__ __forceinline__ void j_diag(T* xr, T* xi, const TileOp& op, int tb0, int tb1) {
    const T r0 = (T)op.m[0], i0 = (T)op.m[1], r1 = (T)op.m[2], i1 = (T)op.m[3];
    const T r2 = (T)op.m[4], i2 = (T)op.m[5], r3 = (T)op.m[6], i3 = (T)op.m[7];
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        if ((s & CINV) != CINV) continue;
        const int b0 = S0 >= 0 ? ((s >> (S0 >= 0 ? S0 : 0)) & 1) : tb0;
        const int b1 = S1 >= 0 ? ((s >> (S1 >= 0 ? S1 : 0)) & 1) : tb1;
        const T pr = b1 ? (b0 ? r3 : r2) : (b0 ? r1 : r0);
        const T pi = b1 ? (b0 ? i3 : i2) : (b0 ? i1 : i0);
        const T yr = pr * xr[s] - pi * xi[s];
        const T yi = pr * xi[s] + pi * xr[s];
        xr[s] = yr;
        xi[s] = yi;
    }
}

// Merged runs of diagonal ops: (pr, pi) *= the op's phase at subcube slot
// ts (1 where the op's in-subcube controls CINV are not all set), and the
// run applied from its 16-entry table (entries agree on the bits outside
// DEP, the subcube positions the run depends on).
// (The phases are read as block-uniform values and selected in registers, as
// in j_uphase_mul.)
template <int S0, int S1, int CINV>
__device__ __forceinline__ void j_phase_mul(double& pr, double& pi, int ts, const TileOp& op) {
    const double r0 = op.m[0], i0 = op.m[1], r1 = op.m[2], i1 = op.m[3];
    const double r2 = op.m[4], i2 = op.m[5], r3 = op.m[6], i3 = op.m[7];
    if ((ts & CINV) != CINV) return;
    const bool b0 = S0 >= 0 && ((ts >> (S0 >= 0 ? S0 : 0)) & 1);
    const bool b1 = S1 >= 0 && ((ts >> (S1 >= 0 ? S1 : 0)) & 1);
    const double dr = b1 ? (b0 ? r3 : r2) : (b0 ? r1 : r0);
    const double di = b1 ? (b0 ? i3 : i2) : (b0 ? i1 : i0);
    const double nr = pr * dr - pi * di;
    pi = pr * di + pi * dr;
    pr = nr;
}

// Diagonal ops that depend only on bits outside the subcube commute with
// every op of the group: (pr, pi) *= phase m[k] when the op applies to the
// thread (its outside controls set); the product is applied once.
// The four phases are read as block-uniform values and the thread's one is
// selected in registers: an index into op.m by the per-thread k would be a
// per-lane memory load in the middle of the group.
__device__ __forceinline__ void j_uphase_mul(double& pr, double& pi, bool on, int k, const TileOp& op) {
    const double r0 = op.m[0], i0 = op.m[1], r1 = op.m[2], i1 = op.m[3];
    const double r2 = op.m[4], i2 = op.m[5], r3 = op.m[6], i3 = op.m[7];
    const bool b0 = (k & 1) != 0, b1 = (k & 2) != 0;
    const double sr = b1 ? (b0 ? r3 : r2) : (b0 ? r1 : r0);
    const double si = b1 ? (b0 ? i3 : i2) : (b0 ? i1 : i0);
    const double dr = on ? sr : 1.0, di = on ? si : 0.0;
    const double nr = pr * dr - pi * di;
    pi = pr * di + pi * dr;
    pr = nr;
}

// A sign op on outside bits only, in the thread phase: -(pr, pi) when on.
__device__ __forceinline__ void j_uphase_neg(double& pr, double& pi, bool on) {
    pr = on ? -pr : pr;
    pi = on ? -pi : pi;
}

// Negate (r, i) when f = 0x80000000 and keep it when f = 0: an XOR of the
// sign bits (the high word's for complex128).
__device__ __forceinline__ void j_flip(double& r, double& i, uint32_t f) {
    const uint64_t m = (uint64_t)f << 32;
    r = __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, r) ^ m);
    i = __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, i) ^ m);
}

__device__ __forceinline__ void j_flip(float& r, float& i, uint32_t f) {
    r = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, r) ^ f);
    i = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, i) ^ f);
}

template <typename T>
__device__ __forceinline__ void j_uphase_apply(T* xr, T* xi, double pr, double pi) {
    const T r = (T)pr, i = (T)pi;
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        const T yr = r * xr[c] - i * xi[c];
        const T yi = r * xi[c] + i * xr[c];
        xr[c] = yr;
        xi[c] = yi;
    }
}

template <int DEP, typename T, typename V>
__device__ __forceinline__ void j_diag_tab(T* xr, T* xi, const V* tab) {
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        const V ph = tab[c & DEP];
        const T yr = ph.x * xr[c] - ph.y * xi[c];
        const T yi = ph.x * xi[c] + ph.y * xr[c];
        xr[c] = yr;
        xi[c] = yi;
    }
}

// ---- pass skeleton of the specialised kernels --------------------------------
// Same tile walk as the generic kernel, with the tile geometry (K tile bits,
// HB of them high) fixed at compile time.
__device__ __forceinline__ uint64_t j_ins0(uint64_t k, int p) {
    return ((k >> p) << (p + 1)) | (k & ((1ull << p) - 1ull));
}

// Global offsets of the 2^HB high-bit patterns of a tile, into LDS.
template <int K, int HB, int NT>
__device__ __forceinline__ void j_pass_begin(uint64_t* hoff, const TilePass& p) {
    for (uint32_t h = threadIdx.x; h < (1u << HB); h += NT) {
        uint64_t o = 0;
        for (int j = 0; j < HB; ++j)
            if ((h >> j) & 1u) o |= 1ull << p.hpos[j];
        hoff[h] = o;
    }
    __syncthreads();
}

// Global index of tile t's element 0.
template <int K, int HB>
__device__ __forceinline__ uint64_t j_tile_base(uint64_t t, const TilePass& p) {
    uint64_t base = t << (K - HB);
#pragma unroll
    for (int j = 0; j < HB; ++j) base = j_ins0(base, p.hpos[j]);
    return base;
}

template <int K, int HB>
__device__ __forceinline__ uint64_t j_gidx(uint64_t base, const uint64_t* hoff, uint32_t e) {
    constexpr int L = K - HB;
    return base | (e & ((1u << L) - 1u)) | hoff[e >> L];
}

// STREAM: nontemporal global accesses, for passes whose states far exceed
// the caches (a value read once per pass gains nothing from being kept).
template <int K, int HB, int NT, bool STREAM = false, typename V>
__device__ __forceinline__ void j_tile_load(V* tile, const V* __restrict__ s, uint64_t base, const uint64_t* hoff) {
    constexpr int EPT = (1 << K) / NT;
    const uint32_t tid = threadIdx.x;
    V r[EPT];
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
        const V* a = s + j_gidx<K, HB>(base, hoff, tid + (uint32_t)j * NT);
        if constexpr (STREAM) r[j] = __builtin_nontemporal_load(a);
        else r[j] = *a;
    }
#pragma unroll
    for (int j = 0; j < EPT; ++j) tile[swz(tid + (uint32_t)j * NT)] = r[j];
    __syncthreads();
}

// The tile of |0...0> instead of a load: element 0 of the tile whose base
// is index 0 is 1, everything else 0 (a circuit's first pass, when the
// state has not been written yet).
template <int K, int NT, typename V>
__device__ __forceinline__ void j_tile_zero(V* tile, uint64_t base) {
    constexpr int EPT = (1 << K) / NT;
    const uint32_t tid = threadIdx.x;
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
        const uint32_t e = tid + (uint32_t)j * NT;
        V v;
        v.x = (base == 0 && e == 0) ? 1 : 0;
        v.y = 0;
        tile[swz(e)] = v;
    }
    __syncthreads();
}

template <int K, int HB, int NT, bool STREAM = false, typename V>
__device__ __forceinline__ void j_tile_store(const V* tile, V* __restrict__ s, uint64_t base, const uint64_t* hoff) {
    constexpr int EPT = (1 << K) / NT;
    const uint32_t tid = threadIdx.x;
    V r[EPT];
#pragma unroll
    for (int j = 0; j < EPT; ++j) r[j] = tile[swz(tid + (uint32_t)j * NT)];
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
        V* a = s + j_gidx<K, HB>(base, hoff, tid + (uint32_t)j * NT);
        if constexpr (STREAM) __builtin_nontemporal_store(r[j], a);
        else *a = r[j];
    }
    __syncthreads();
}

// One group: every thread gathers its 16-amplitude subcube over tile bits
// S0..S3 into registers, runs body(xr, xi, gb) (gb: the subcube's global
// base index) and scatters it back.  gflip(gb) / sflip(gb): tile bits by
// which the thread's gather / scatter slots are flipped (X gates whose
// control lies outside the subcube, applied by addressing when they are the
// first / last op of the group on their bit).
template <int K, int HB, int NT, int S0, int S1, int S2, int S3, typename T, typename V, typename F, typename GF,
          typename SF>
__device__ __forceinline__ void j_group(V* tile, uint64_t base, const uint64_t* hoff, F&& body, GF&& gflip,
                                        SF&& sflip) {
    constexpr uint32_t NSUB = (1u << K) >> 4;
    constexpr bool SFULL = NSUB >= (uint32_t)NT;
    constexpr int SPT = SFULL ? (int)(NSUB / NT) : 1;
    uint32_t tid = threadIdx.x;
    // Opaque to the optimiser, so the slot addresses below are formed in
    // every tile iteration: hoisted out of the tile loop, every group's 16
    // addresses would stay live across the whole pass and spill.
    asm volatile("" : "+v"(tid));
#pragma unroll
    for (int k = 0; k < SPT; ++k) {
        const uint32_t q = tid + (uint32_t)k * NT;
        if (!SFULL && q >= NSUB) break;
        const uint32_t eb = ins0(ins0(ins0(ins0(q, S0), S1), S2), S3);
        const uint64_t gb = j_gidx<K, HB>(base, hoff, eb);
        // swz is linear over XOR and the slot bits are disjoint from eb's:
        // slot c sits at swz(eb) ^ swz(offset of c), one XOR per slot
        const uint32_t sb = swz(eb);
        const uint32_t sg = sb ^ swz(gflip(gb));
        T xr[16], xi[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            const uint32_t oc = ((uint32_t)(c & 1) << S0) | ((uint32_t)((c >> 1) & 1) << S1) |
                                ((uint32_t)((c >> 2) & 1) << S2) | ((uint32_t)((c >> 3) & 1) << S3);
            const V v = tile[sg ^ swz(oc)];
            xr[c] = v.x;
            xi[c] = v.y;
        }
        body(xr, xi, gb);
        const uint32_t ss = sb ^ swz(sflip(gb));
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            const uint32_t oc = ((uint32_t)(c & 1) << S0) | ((uint32_t)((c >> 1) & 1) << S1) |
                                ((uint32_t)((c >> 2) & 1) << S2) | ((uint32_t)((c >> 3) & 1) << S3);
            V v;
            v.x = xr[c];
            v.y = xi[c];
            tile[ss ^ swz(oc)] = v;
        }
    }
    __syncthreads();
}

// Global offset of subcube slot c relative to slot 0: the tile's element ->
// index map is a bit spread (linear over XOR), so slot c of a subcube at
// global index g sits at g ^ d(c), d(c) the XOR of the spread subcube bits.
template <int K, int HB, int S>
__device__ __forceinline__ uint64_t j_spread_bit(const TilePass& p) {
    constexpr int L = K - HB;
    if constexpr (S < L) return 1ull << S;
    else return 1ull << p.hpos[S - L];
}

// The first group of a tile, straight from the state: each thread gathers
// its subcube's 16 amplitudes from global memory (INIT: from |0...0>, the
// state not yet written), applies the group, and, unless LAST, scatters into
// the LDS tile after a barrier that retires the previous tile's LDS reads.
// LAST (a pass of one group): the subcube goes straight back to the state
// and the pass touches no LDS.  Coalesced when the subcube bits leave the
// lowest tile bits to the lanes (whole 128-byte lines per wave access).
template <int K, int HB, int NT, int S0, int S1, int S2, int S3, bool STREAM, bool INIT, bool LAST, typename T,
          typename V, typename F, typename GF, typename SF>
__device__ __forceinline__ void j_group_first(V* tile, V* __restrict__ st, uint64_t base, const uint64_t* hoff,
                                              const TilePass& p, F&& body, GF&& gflip, SF&& sflip) {
    constexpr uint32_t NSUB = (1u << K) >> 4;
    static_assert(NSUB <= (uint32_t)NT, "one subcube per thread at most");
    uint32_t tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const bool act = NSUB == (uint32_t)NT || tid < NSUB;
    const uint64_t d0 = j_spread_bit<K, HB, S0>(p), d1 = j_spread_bit<K, HB, S1>(p);
    const uint64_t d2 = j_spread_bit<K, HB, S2>(p), d3 = j_spread_bit<K, HB, S3>(p);
    T xr[16], xi[16];
    uint32_t eb = 0;
    uint64_t gb = 0;
    if (act) {
        eb = ins0(ins0(ins0(ins0(tid, S0), S1), S2), S3);
        gb = j_gidx<K, HB>(base, hoff, eb);
        const uint64_t g0 = j_gidx<K, HB>(base, hoff, eb ^ gflip(gb));
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            const uint64_t gi = g0 ^ ((c & 1) ? d0 : 0ull) ^ ((c & 2) ? d1 : 0ull) ^ ((c & 4) ? d2 : 0ull) ^
                                ((c & 8) ? d3 : 0ull);
            V v;
            if constexpr (INIT) {
                v.x = gi == 0 ? 1 : 0;
                v.y = 0;
            } else if constexpr (STREAM) {
                v = __builtin_nontemporal_load(st + gi);
            } else {
                v = st[gi];
            }
            xr[c] = v.x;
            xi[c] = v.y;
        }
        body(xr, xi, gb);
    }
    if constexpr (LAST) {
        if (act) {
            const uint64_t g0 = j_gidx<K, HB>(base, hoff, eb ^ sflip(gb));
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                const uint64_t gi = g0 ^ ((c & 1) ? d0 : 0ull) ^ ((c & 2) ? d1 : 0ull) ^ ((c & 4) ? d2 : 0ull) ^
                                    ((c & 8) ? d3 : 0ull);
                V v;
                v.x = xr[c];
                v.y = xi[c];
                if constexpr (STREAM) __builtin_nontemporal_store(v, st + gi);
                else st[gi] = v;
            }
        }
    } else {
        __syncthreads();
        if (act) {
            const uint32_t ss = swz(eb) ^ swz(sflip(gb));
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                const uint32_t oc = ((uint32_t)(c & 1) << S0) | ((uint32_t)((c >> 1) & 1) << S1) |
                                    ((uint32_t)((c >> 2) & 1) << S2) | ((uint32_t)((c >> 3) & 1) << S3);
                V v;
                v.x = xr[c];
                v.y = xi[c];
                tile[ss ^ swz(oc)] = v;
            }
        }
        __syncthreads();
    }
}

// The first group split in two so that a tile's gather can be issued one
// tile ahead (software pipelining across the tiles of a block):
// j_first_fetch loads the subcube of the tile at `base` into r[16] without
// waiting for it; j_first_apply takes r into registers, calls `next` (which
// issues the next tile's j_first_fetch into r: those loads go out before this
// tile's stores, so waiting for them later never waits for the stores),
// applies the group and scatters it (into the LDS tile after a barrier, or,
// LAST, straight back to the state).
template <int K, int HB, int NT, int S0, int S1, int S2, int S3, bool STREAM, typename V, typename GF>
__device__ __forceinline__ void j_first_fetch(V* r, const V* __restrict__ st, uint64_t base, const uint64_t* hoff,
                                              const TilePass& p, GF&& gflip) {
    constexpr uint32_t NSUB = (1u << K) >> 4;
    static_assert(NSUB <= (uint32_t)NT, "one subcube per thread at most");
    uint32_t tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    if (NSUB == (uint32_t)NT || tid < NSUB) {
        const uint64_t d0 = j_spread_bit<K, HB, S0>(p), d1 = j_spread_bit<K, HB, S1>(p);
        const uint64_t d2 = j_spread_bit<K, HB, S2>(p), d3 = j_spread_bit<K, HB, S3>(p);
        const uint32_t eb = ins0(ins0(ins0(ins0(tid, S0), S1), S2), S3);
        const uint64_t gb = j_gidx<K, HB>(base, hoff, eb);
        const uint64_t g0 = j_gidx<K, HB>(base, hoff, eb ^ gflip(gb));
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            const uint64_t gi = g0 ^ ((c & 1) ? d0 : 0ull) ^ ((c & 2) ? d1 : 0ull) ^ ((c & 4) ? d2 : 0ull) ^
                                ((c & 8) ? d3 : 0ull);
            if constexpr (STREAM) r[c] = __builtin_nontemporal_load(st + gi);
            else r[c] = st[gi];
        }
    }
}

template <int K, int HB, int NT, int S0, int S1, int S2, int S3, bool STREAM, bool LAST, bool EARLY, typename T,
          typename V, typename F, typename SF, typename NX>
__device__ __forceinline__ void j_first_apply(V* tile, V* __restrict__ st, V* r, uint64_t base,
                                              const uint64_t* hoff, const TilePass& p, F&& body, SF&& sflip,
                                              NX&& next) {
    constexpr uint32_t NSUB = (1u << K) >> 4;
    uint32_t tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const bool act = NSUB == (uint32_t)NT || tid < NSUB;
    T xr[16], xi[16];
    uint32_t eb = 0;
    uint64_t gb = 0;
    if (act) {
        eb = ins0(ins0(ins0(ins0(tid, S0), S1), S2), S3);
        gb = j_gidx<K, HB>(base, hoff, eb);
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            xr[c] = r[c].x;
            xi[c] = r[c].y;
        }
    }
    if constexpr (EARLY) next();
    if (act) body(xr, xi, gb);
    if constexpr (LAST) {
        if (act) {
            const uint64_t d0 = j_spread_bit<K, HB, S0>(p), d1 = j_spread_bit<K, HB, S1>(p);
            const uint64_t d2 = j_spread_bit<K, HB, S2>(p), d3 = j_spread_bit<K, HB, S3>(p);
            const uint64_t g0 = j_gidx<K, HB>(base, hoff, eb ^ sflip(gb));
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                const uint64_t gi = g0 ^ ((c & 1) ? d0 : 0ull) ^ ((c & 2) ? d1 : 0ull) ^ ((c & 4) ? d2 : 0ull) ^
                                    ((c & 8) ? d3 : 0ull);
                V v;
                v.x = xr[c];
                v.y = xi[c];
                if constexpr (STREAM) __builtin_nontemporal_store(v, st + gi);
                else st[gi] = v;
            }
        }
    } else {
        __syncthreads();
        if (act) {
            const uint32_t ss = swz(eb) ^ swz(sflip(gb));
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                const uint32_t oc = ((uint32_t)(c & 1) << S0) | ((uint32_t)((c >> 1) & 1) << S1) |
                                    ((uint32_t)((c >> 2) & 1) << S2) | ((uint32_t)((c >> 3) & 1) << S3);
                V v;
                v.x = xr[c];
                v.y = xi[c];
                tile[ss ^ swz(oc)] = v;
            }
        }
        __syncthreads();
    }
    if constexpr (!EARLY) next();
}

// The last group of a tile, straight to the state: gathers from the LDS tile,
// applies the group and stores the subcube to global memory.  BAR: a closing
// barrier (needed when the next tile's first phase writes the LDS tile
// without one of its own).
template <int K, int HB, int NT, int S0, int S1, int S2, int S3, bool STREAM, bool BAR, typename T, typename V,
          typename F, typename GF, typename SF>
__device__ __forceinline__ void j_group_last(const V* tile, V* __restrict__ st, uint64_t base, const uint64_t* hoff,
                                             const TilePass& p, F&& body, GF&& gflip, SF&& sflip) {
    constexpr uint32_t NSUB = (1u << K) >> 4;
    static_assert(NSUB <= (uint32_t)NT, "one subcube per thread at most");
    uint32_t tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    if (NSUB == (uint32_t)NT || tid < NSUB) {
        const uint64_t d0 = j_spread_bit<K, HB, S0>(p), d1 = j_spread_bit<K, HB, S1>(p);
        const uint64_t d2 = j_spread_bit<K, HB, S2>(p), d3 = j_spread_bit<K, HB, S3>(p);
        const uint32_t eb = ins0(ins0(ins0(ins0(tid, S0), S1), S2), S3);
        const uint64_t gb = j_gidx<K, HB>(base, hoff, eb);
        const uint32_t sg = swz(eb) ^ swz(gflip(gb));
        T xr[16], xi[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            const uint32_t oc = ((uint32_t)(c & 1) << S0) | ((uint32_t)((c >> 1) & 1) << S1) |
                                ((uint32_t)((c >> 2) & 1) << S2) | ((uint32_t)((c >> 3) & 1) << S3);
            const V v = tile[sg ^ swz(oc)];
            xr[c] = v.x;
            xi[c] = v.y;
        }
        body(xr, xi, gb);
        const uint64_t g0 = j_gidx<K, HB>(base, hoff, eb ^ sflip(gb));
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            const uint64_t gi = g0 ^ ((c & 1) ? d0 : 0ull) ^ ((c & 2) ? d1 : 0ull) ^ ((c & 4) ? d2 : 0ull) ^
                                ((c & 8) ? d3 : 0ull);
            V v;
            v.x = xr[c];
            v.y = xi[c];
            if constexpr (STREAM) __builtin_nontemporal_store(v, st + gi);
            else st[gi] = v;
        }
    }
    if constexpr (BAR) __syncthreads();
}

}  // namespace lq
