// Several flip groups of a Pauli sum from one read of the state.
//
// The flip masks x_0 .. x_{m-1} of m groups (linearly independent over
// GF(2)) span a space X of 2^m masks.  The state splits into orbits
// {i0 ^ X[c] : c < 2^m}, X[c] = XOR of the x_j with bit j set in c, and an
// orbit's base index i0 has zeros at m pivot bits (bits onto which X
// projects one to one, found by Gaussian elimination).  A thread loads the
// 2^m amplitudes of an orbit once and evaluates every group of the item on
// them: group j pairs element c with c ^ (1 << j).  For a pair (i, j = i ^ x)
// and a term t with sign s_t(i) = (-1)^{popc(i & zy_t)}, the two rows of the
// term (PauliTerm::expectation, src/optimization/observable.rs:95-147) add up
// to s_t(i) * g_t * Re(conj(psi_i) psi_j) when popc(x & zy_t) is even
// (g_t = 2 cr_t) and s_t(i) * g_t * Im(conj(psi_i) psi_j) when it is odd
// (g_t = -2 ci_t); the identity holds with either element as i.  With i the
// orbit's element c, s_t(i0 ^ X[c]) = s_t(i0) * sigma_t(c), and sigma_t is a
// per-term bit table over c built on the host.  Diagonal terms (x = 0) add
// cr_t * s_t(i) * |psi_i|^2 over the orbit's elements.
//
// m groups cost one read of the state instead of m, and each amplitude is
// loaded by one lane of a wavefront whose 64 lanes take 64 consecutive base
// indices, so every load of a wavefront is one permuted run of consecutive
// amplitudes.
#pragma once

#include <cstdint>

namespace lq {

constexpr int kOrbitMax = 4;  // up to 16 amplitudes per thread

struct OrbitItem {
    uint64_t state_off;  // amplitude offset of the item's state
    uint64_t work;       // orbits: dim >> m
    int32_t m;           // groups (generators); 0: diagonal terms only
    int32_t piv[kOrbitMax];  // pivot bits, ascending
    int32_t x_first;     // X[0 .. 2^m) at xtab + x_first
    int32_t term_first;  // OrbitTerm records
    int32_t term_count;
    int32_t block_first; // first block of the item in the launch
    int32_t nblocks;
    int32_t inst;
};

struct OrbitTerm {
    uint64_t zy;     // bits carrying Z or Y
    double g;        // flip: 2 cr or -2 ci; diagonal: cr
    int32_t gen;     // the term's group within the item; -1: diagonal
    int32_t imag;    // flip: the Im part of conj(psi_i) psi_j
    uint32_t sigma;  // bit c: popc(X[c] & zy) odd
    int32_t pad;
};

#ifdef __HIPCC__
// Sum over the orbits k0, k0 + stride, ... < it.work of the item's terms.
template <int M, typename V>
__device__ __forceinline__ double orbit_acc(const V* __restrict__ s, const OrbitItem& it,
                                            const uint64_t* __restrict__ xtab, const OrbitTerm* __restrict__ tm,
                                            uint64_t k0, uint64_t stride) {
    constexpr int C = 1 << M;
    // the masks are the same for the whole block: scalar registers
    uint64_t xc[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const uint64_t v = xtab[it.x_first + c];
        xc[c] = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32)) << 32) |
                __builtin_amdgcn_readfirstlane((uint32_t)v);
    }
    int piv[M > 0 ? M : 1];
#pragma unroll
    for (int j = 0; j < M; ++j) piv[j] = it.piv[j];
    double acc = 0.0;
    for (uint64_t k = k0; k < it.work; k += stride) {
        uint64_t i0 = k;
#pragma unroll
        for (int j = 0; j < M; ++j) i0 = ((i0 >> piv[j]) << (piv[j] + 1)) | (i0 & ((1ull << piv[j]) - 1ull));
        V a[C];
#pragma unroll
        for (int c = 0; c < C; ++c) a[c] = s[i0 ^ xc[c]];
        // terms come ordered by group (plan_orbits: generators 0 .. M-1,
        // then the diagonal terms): each group's pair products are formed
        // once and kept only while its terms are summed
        int t = it.term_first;
        const int t1 = it.term_first + it.term_count;
#pragma unroll
        for (int J = 0; J < M; ++J) {
            if (t < t1 && tm[t].gen == J) {
                double zr[C / 2 > 0 ? C / 2 : 1], zi[C / 2 > 0 ? C / 2 : 1];
                int np = 0;
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    if (c & (1 << J)) continue;
                    const V u = a[c], v = a[c | (1 << J)];
                    zr[np] = (double)u.x * (double)v.x + (double)u.y * (double)v.y;
                    zi[np] = (double)u.x * (double)v.y - (double)u.y * (double)v.x;
                    ++np;
                }
                do {
                    const OrbitTerm q = tm[t];
                    const double g = (__popcll(i0 & q.zy) & 1) ? -q.g : q.g;
                    double wr = 0.0, wi = 0.0;
                    int e = 0;
#pragma unroll
                    for (int c = 0; c < C; ++c) {
                        if (c & (1 << J)) continue;
                        const bool neg = (q.sigma >> c) & 1u;
                        wr += neg ? -zr[e] : zr[e];
                        wi += neg ? -zi[e] : zi[e];
                        ++e;
                    }
                    acc += g * (q.imag ? wi : wr);
                    ++t;
                } while (t < t1 && tm[t].gen == J);
            }
        }
        // then the diagonal terms
        for (; t < t1; ++t) {
            const OrbitTerm q = tm[t];
            const double g = (__popcll(i0 & q.zy) & 1) ? -q.g : q.g;
            double w = 0.0;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const double p = (double)a[c].x * (double)a[c].x + (double)a[c].y * (double)a[c].y;
                w += ((q.sigma >> c) & 1u) ? -p : p;
            }
            acc += g * w;
        }
    }
    return acc;
}

constexpr int kOrbitThreads = 256;

// One launch for every item: the block's item by binary search over
// block_first, its share of the item's orbits, one partial sum per block.
template <typename V>
__global__ void __launch_bounds__(kOrbitThreads) k_expect_orbits(const V* __restrict__ psi,
                                                                 const OrbitItem* __restrict__ items, int nitems,
                                                                 const uint64_t* __restrict__ xtab,
                                                                 const OrbitTerm* __restrict__ terms,
                                                                 double* __restrict__ partial) {
    __shared__ double sh[kOrbitThreads];
    int lo = 0, hi = nitems - 1;
    const int b = blockIdx.x;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].block_first <= b) lo = mid;
        else hi = mid - 1;
    }
    const OrbitItem it = items[lo];
    const V* s = psi + it.state_off;
    const uint64_t k0 = (uint64_t)(b - it.block_first) * kOrbitThreads + threadIdx.x;
    const uint64_t stride = (uint64_t)it.nblocks * kOrbitThreads;
    double acc = 0.0;
    switch (it.m) {
        case 0: acc = orbit_acc<0, V>(s, it, xtab, terms, k0, stride); break;
        case 1: acc = orbit_acc<1, V>(s, it, xtab, terms, k0, stride); break;
        case 2: acc = orbit_acc<2, V>(s, it, xtab, terms, k0, stride); break;
        case 3: acc = orbit_acc<3, V>(s, it, xtab, terms, k0, stride); break;
        default: acc = orbit_acc<4, V>(s, it, xtab, terms, k0, stride); break;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int k = kOrbitThreads / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[b] = sh[0];
}
#endif

}  // namespace lq
