// Pauli-sum expectation values: the replacement of PauliTerm::expectation /
// PauliObservable::expectation (src/optimization/observable.rs:95-147,
// 228-231).
//
// For a term with flip mask x (bits carrying X or Y) and sign mask zy (bits
// carrying Z or Y) the reference evaluates
//     Re sum_i conj(psi_i) * i^{#Y} * (-1)^{popcount(i & zy)} * psi_{i ^ x}
// (the Y phase comes from the row bit, observable.rs:125-133).  Terms sharing
// a flip mask share the products conj(psi_i) psi_{i^x}, so they are grouped
// and each group costs one read of the state: the kernel walks the pairs
// {i, i ^ x} once (i with the top bit of x clear) and adds the contributions
// of both rows.  Accumulation is f64 for both amplitude types, in a fixed
// grid/tree order (bit-reproducible).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>

#include "lq_host.h"

namespace lq {

constexpr int kExpBlocksMax = 2048;
constexpr int kExpThreads = 256;

template <int NT>
__device__ __forceinline__ double block_sum_e(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    return sh[0];
}

// x == 0: diagonal group, sum_i |psi_i|^2 * w(i), w(i) = sum_t cr_t
// (-1)^{popc(i & zy_t)}.  A wavefront reads 64 consecutive amplitudes (one
// chunk, i = chunk * 64 + lane) and walks R consecutive chunks per work item
// (R = 64, fewer for small states), so bits 6..6+log2(R) of i are the step r
// and the rest are fixed for the item.  Terms whose mask misses the r bits
// ("static": for a local Hamiltonian, most of them) are summed once per item
// and lane into w_static; the others ("dynamic", at most kDiagDyn per launch)
// keep their fixed-bit sign in d_t, and their r-sign comes from a 64-bit
// table per term (bit r = parity of r & the term's r bits) on the scalar
// unit, so a chunk costs one FMA per dynamic term.  The host orders each
// launch's terms static first.
constexpr int kDiagDyn = 16;
template <typename V>
__global__ void __launch_bounds__(kExpThreads) k_expect_diag(const V* __restrict__ psi, uint64_t items, uint32_t R,
                                                             uint32_t lane_limit,
                                                             const PauliTermDev* __restrict__ terms, int ns, int nd,
                                                             double* __restrict__ partial) {
    __shared__ double sh[kExpThreads];
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t w0 = (uint64_t)__builtin_amdgcn_readfirstlane(blockIdx.x * (kExpThreads / 64) +
                                                                 (threadIdx.x >> 6));
    const uint64_t W = (uint64_t)gridDim.x * (kExpThreads / 64);
    // r-sign tables of the dynamic terms (uniform)
    uint64_t rsign[kDiagDyn];
#pragma unroll
    for (int t = 0; t < kDiagDyn; ++t) {
        uint64_t m = 0;
        if (t < nd) {
            const uint64_t zr = (terms[ns + t].zy >> 6) & (uint64_t)(R - 1);
            for (uint32_t r = 0; r < R; ++r) m |= (uint64_t)(__popcll(r & zr) & 1) << r;
        }
        rsign[t] = m;
    }
    constexpr uint32_t U = 4;  // chunks in flight per wave
    double acc = 0.0;
    if (lane < lane_limit) {
        for (uint64_t it = w0; it < items; it += W) {
            const uint64_t base = ((it * R) << 6) | lane;
            double wst = 0.0;
            for (int t = 0; t < ns; ++t) wst += (__popcll(base & terms[t].zy) & 1) ? -terms[t].cr : terms[t].cr;
            double d[kDiagDyn];
#pragma unroll
            for (int t = 0; t < kDiagDyn; ++t) {
                d[t] = 0.0;
                if (t < nd) d[t] = (__popcll(base & terms[ns + t].zy) & 1) ? -terms[ns + t].cr : terms[ns + t].cr;
            }
            for (uint32_t r = 0; r < R; r += U) {
                V a[U];
#pragma unroll
                for (uint32_t j = 0; j < U; ++j)
                    if (r + j < R) a[j] = psi[base | ((uint64_t)(r + j) << 6)];
#pragma unroll
                for (uint32_t j = 0; j < U; ++j) {
                    if (r + j >= R) break;
                    const double p = (double)a[j].x * (double)a[j].x + (double)a[j].y * (double)a[j].y;
                    double w = wst;
#pragma unroll
                    for (int t = 0; t < kDiagDyn; ++t)
                        if (t < nd) w = __builtin_fma(((rsign[t] >> (r + j)) & 1) ? -1.0 : 1.0, d[t], w);
                    acc += p * w;
                }
            }
        }
    }
    double s = block_sum_e<kExpThreads>(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// x != 0: pairs {i, j = i ^ x} with bit h (top bit of x) of i clear.  A
// wavefront takes 64 consecutive pair indices k (i = k with a zero inserted
// at bit h), U chunks of them in flight.
template <typename V>
__global__ void __launch_bounds__(kExpThreads) k_expect_flip(const V* __restrict__ psi, uint64_t chunks,
                                                             uint32_t lane_limit, uint64_t x, int h,
                                                             const PauliTermDev* __restrict__ terms, int nt,
                                                             double* __restrict__ partial) {
    __shared__ double sh[kExpThreads];
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t w0 = (uint64_t)__builtin_amdgcn_readfirstlane(blockIdx.x * (kExpThreads / 64) +
                                                                 (threadIdx.x >> 6));
    const uint64_t W = (uint64_t)gridDim.x * (kExpThreads / 64);
    constexpr int U = 4;
    double acc = 0.0;
    if (lane < lane_limit) {
        for (uint64_t c = w0; c < chunks; c += U * W) {
            V a[U], b[U];
            uint64_t ii[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                ii[u] = insert_zero(((c + u * W) << 6) | lane, h);
                if (c + u * W < chunks) {
                    a[u] = psi[ii[u]];
                    b[u] = psi[ii[u] ^ x];
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (c + u * W >= chunks) break;
                const uint64_t i = ii[u], j = i ^ x;
                // z = conj(a) * b
                const double zr = (double)a[u].x * (double)b[u].x + (double)a[u].y * (double)b[u].y;
                const double zi = (double)a[u].x * (double)b[u].y - (double)a[u].y * (double)b[u].x;
                double wir = 0.0, wii = 0.0, wjr = 0.0, wji = 0.0;
                for (int t = 0; t < nt; ++t) {
                    const PauliTermDev tm = terms[t];
                    const bool si = __popcll(i & tm.zy) & 1;
                    const bool sj = __popcll(j & tm.zy) & 1;
                    wir += si ? -tm.cr : tm.cr;
                    wii += si ? -tm.ci : tm.ci;
                    wjr += sj ? -tm.cr : tm.cr;
                    wji += sj ? -tm.ci : tm.ci;
                }
                // Re(z * w_i) + Re(conj(z) * w_j)
                acc += zr * wir - zi * wii + zr * wjr + zi * wji;
            }
        }
    }
    double s = block_sum_e<kExpThreads>(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void __launch_bounds__(1024) k_expect_final(const double* __restrict__ partial, int count,
                                                      double* __restrict__ out) {
    __shared__ double sh[1024];
    double s = 0.0;
    for (int i = threadIdx.x; i < count; i += blockDim.x) s += partial[i];
    double t = block_sum_e<1024>(s, sh);
    if (threadIdx.x == 0) out[0] = t;
}

int compile_pauli_sum(int n, const uint8_t* paulis, const double* coeffs, int num_terms, int convention,
                      PauliSum* out) {
    if (num_terms < 0 || (num_terms > 0 && (!paulis || !coeffs))) {
        set_error("invalid Pauli sum");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    if (convention != LQ_PAULI_REFERENCE && convention != LQ_PAULI_PHYSICAL) {
        set_error("unknown Pauli convention %d", convention);
        return LQ_ERR_INVALID_ARGUMENT;
    }
    out->n = n;
    out->terms.clear();
    out->groups.clear();
    // group by flip mask, keeping first-appearance order
    std::vector<uint64_t> xs;
    std::vector<std::vector<PauliTermDev>> per;
    std::map<uint64_t, int> where;
    for (int t = 0; t < num_terms; ++t) {
        uint64_t x = 0, zy = 0;
        int ny = 0;
        for (int q = 0; q < n; ++q) {
            uint8_t p = paulis[(size_t)t * n + q];
            uint64_t b = 1ull << (n - 1 - q);
            if (p > 3) {
                set_error("invalid Pauli code %d", (int)p);
                return LQ_ERR_INVALID_ARGUMENT;
            }
            if (p == LQ_PAULI_X || p == LQ_PAULI_Y) x |= b;
            if (p == LQ_PAULI_Z || p == LQ_PAULI_Y) zy |= b;
            if (p == LQ_PAULI_Y) ny++;
        }
        double c = coeffs[t];
        int k = ny & 3;
        if (convention == LQ_PAULI_PHYSICAL) k = (4 - k) & 3;  // (-i)^{#Y}
        PauliTermDev d;
        d.zy = zy;
        d.cr = (k == 0) ? c : (k == 2) ? -c : 0.0;
        d.ci = (k == 1) ? c : (k == 3) ? -c : 0.0;
        auto it = where.find(x);
        if (it == where.end()) {
            where[x] = (int)xs.size();
            xs.push_back(x);
            per.emplace_back();
            per.back().push_back(d);
        } else {
            per[it->second].push_back(d);
        }
    }
    for (size_t g = 0; g < xs.size(); ++g) {
        PauliGroup grp;
        grp.x = xs[g];
        grp.first = (int)out->terms.size();
        grp.count = (int)per[g].size();
        out->groups.push_back(grp);
        out->terms.insert(out->terms.end(), per[g].begin(), per[g].end());
    }
    return LQ_OK;
}

template <typename V>
static int expect_t(const lq_state* st, const PauliSum& ps, double* out) {
    const int ng = (int)ps.groups.size();
    if (ng == 0) {
        *out = 0.0;
        return LQ_OK;
    }
    // device layout: [terms][partials...][result]
    size_t term_bytes = ps.terms.size() * sizeof(PauliTermDev);
    size_t term_pad = (term_bytes + 255) & ~(size_t)255;
    const uint64_t chunks = st->dim >= 64 ? st->dim >> 6 : 1;
    const uint32_t lane_limit = st->dim >= 64 ? 64u : (uint32_t)st->dim;
    const uint32_t R = (uint32_t)std::min<uint64_t>(chunks, 64);
    const uint64_t items = chunks / R;
    const uint64_t rmask = (uint64_t)(R - 1) << 6;
    int nb_diag = (int)std::min<uint64_t>(std::max<uint64_t>((items + 3) / 4, 1), kExpBlocksMax);
    const uint64_t half = st->dim >> 1;
    const uint64_t fchunks = half >= 64 ? half >> 6 : 1;
    const uint32_t flanes = half >= 64 ? 64u : (uint32_t)half;
    int nb_flip = (int)std::min<uint64_t>(std::max<uint64_t>((fchunks + 15) / 16, 1), kExpBlocksMax);
    int nb = std::max(nb_diag, nb_flip);
    // Device term table: the diagonal group's terms are reordered static
    // first (masks that miss the r bits), then dynamic; a launch takes every
    // static term and up to kDiagDyn dynamic ones, further launches the rest.
    std::vector<PauliTermDev> dev = ps.terms;
    int diag_first = -1, diag_static = 0, diag_count = 0;
    size_t launches = 0;
    for (const PauliGroup& grp : ps.groups) {
        if (grp.x != 0) {
            ++launches;
            continue;
        }
        std::vector<PauliTermDev> st_, dy_;
        for (int t = grp.first; t < grp.first + grp.count; ++t)
            ((ps.terms[t].zy & rmask) == 0 ? st_ : dy_).push_back(ps.terms[t]);
        std::copy(st_.begin(), st_.end(), dev.begin() + grp.first);
        std::copy(dy_.begin(), dy_.end(), dev.begin() + grp.first + st_.size());
        diag_first = grp.first;
        diag_static = (int)st_.size();
        diag_count = grp.count;
        launches += std::max<size_t>(1, (dy_.size() + kDiagDyn - 1) / kDiagDyn);
    }
    size_t part_count = launches * nb;
    char* base = (char*)scratch(term_pad + (part_count + 8) * sizeof(double));
    if (!base) {
        set_error("scratch allocation failed");
        return LQ_ERR_OUT_OF_MEMORY;
    }
    PauliTermDev* d_terms = (PauliTermDev*)base;
    double* d_part = (double*)(base + term_pad);
    double* d_out = d_part + part_count;
    double* h = host_staging(term_bytes / sizeof(double) + 16);
    if (!h) return LQ_ERR_OUT_OF_MEMORY;
    std::memcpy(h, dev.data(), term_bytes);
    LQ_HIP_CHECK(hipMemcpyAsync(d_terms, h, term_bytes, hipMemcpyHostToDevice, stream()));
    int written = 0;
    for (int g = 0; g < ng; ++g) {
        const PauliGroup& grp = ps.groups[g];
        if (grp.x == 0) {
            // launch 0: the static terms and the first dynamic ones
            const int nd_all = diag_count - diag_static;
            int t0 = 0;
            do {
                const int ns = t0 == 0 ? diag_static : 0;
                const int nd = std::min(kDiagDyn, nd_all - t0);
                const int first = diag_first + (t0 == 0 ? 0 : diag_static + t0);
                hipLaunchKernelGGL(k_expect_diag<V>, dim3(nb_diag), dim3(kExpThreads), 0, stream(),
                                   (const V*)st->d_amps, items, R, lane_limit,
                                   (const PauliTermDev*)(d_terms + first), ns, std::max(nd, 0), d_part + written);
                written += nb_diag;
                t0 += kDiagDyn;
            } while (t0 < nd_all);
        } else {
            int hbit = 63 - __builtin_clzll(grp.x);
            hipLaunchKernelGGL(k_expect_flip<V>, dim3(nb_flip), dim3(kExpThreads), 0, stream(),
                               (const V*)st->d_amps, fchunks, flanes, grp.x, hbit,
                               (const PauliTermDev*)(d_terms + grp.first), grp.count, d_part + written);
            written += nb_flip;
        }
        LQ_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_expect_final, dim3(1), dim3(1024), 0, stream(), (const double*)d_part, written, d_out);
    LQ_LAUNCH_CHECK();
    // the staging buffer is reused for the result only after the term copy
    // has completed (same stream, synchronised below)
    LQ_HIP_CHECK(hipStreamSynchronize(stream()));
    LQ_HIP_CHECK(hipMemcpyAsync(h, d_out, sizeof(double), hipMemcpyDeviceToHost, stream()));
    LQ_HIP_CHECK(hipStreamSynchronize(stream()));
    *out = h[0];
    return LQ_OK;
}

int expect_pauli_sum(const lq_state* st, const PauliSum& ps, double* out) {
    if (ps.n != st->n) {
        // "Pauli term must have operator for each qubit" (observable.rs:97-101)
        set_error("Pauli term must have operator for each qubit (%d terms' qubits vs state %d)", ps.n, st->n);
        return LQ_ERR_DIMENSION_MISMATCH;
    }
    if (st->dtype == LQ_C128) return expect_t<double2>(st, ps, out);
    return expect_t<float2>(st, ps, out);
}

}  // namespace lq

using namespace lq;

extern "C" int lq_expect_pauli_sum(const lq_state* st, const uint8_t* paulis, const double* coeffs, int num_terms,
                                   int convention, double* out) {
    clear_error();
    int rc = check_state(st);
    if (rc) return rc;
    if (!out) {
        set_error("out is null");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    PauliSum ps;
    rc = compile_pauli_sum(st->n, paulis, coeffs, num_terms, convention, &ps);
    if (rc) return rc;
    return expect_pauli_sum(st, ps, out);
}
