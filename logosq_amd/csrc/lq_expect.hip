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

// x == 0: diagonal group, sum_i |psi_i|^2 * sum_t cr_t (-1)^{popc(i & zy_t)}.
template <typename V>
__global__ void __launch_bounds__(kExpThreads) k_expect_diag(const V* __restrict__ psi, uint64_t dim,
                                                             const PauliTermDev* __restrict__ terms, int nt,
                                                             double* __restrict__ partial) {
    __shared__ double sh[kExpThreads];
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    double acc = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
        V a = psi[i];
        double p = (double)a.x * (double)a.x + (double)a.y * (double)a.y;
        double w = 0.0;
        for (int t = 0; t < nt; ++t) {
            double c = terms[t].cr;
            w += (__popcll(i & terms[t].zy) & 1) ? -c : c;
        }
        acc += p * w;
    }
    double s = block_sum_e<kExpThreads>(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// x != 0: pairs {i, j = i ^ x} with bit h (top bit of x) of i clear.
template <typename V>
__global__ void __launch_bounds__(kExpThreads) k_expect_flip(const V* __restrict__ psi, uint64_t half, uint64_t x,
                                                             int h, const PauliTermDev* __restrict__ terms, int nt,
                                                             double* __restrict__ partial) {
    __shared__ double sh[kExpThreads];
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    double acc = 0.0;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < half; k += stride) {
        uint64_t i = insert_zero(k, h);
        uint64_t j = i ^ x;
        V a = psi[i], b = psi[j];
        // z = conj(a) * b
        double zr = (double)a.x * (double)b.x + (double)a.y * (double)b.y;
        double zi = (double)a.x * (double)b.y - (double)a.y * (double)b.x;
        double wir = 0.0, wii = 0.0, wjr = 0.0, wji = 0.0;
        for (int t = 0; t < nt; ++t) {
            PauliTermDev tm = terms[t];
            bool si = __popcll(i & tm.zy) & 1;
            bool sj = __popcll(j & tm.zy) & 1;
            wir += si ? -tm.cr : tm.cr;
            wii += si ? -tm.ci : tm.ci;
            wjr += sj ? -tm.cr : tm.cr;
            wji += sj ? -tm.ci : tm.ci;
        }
        // Re(z * w_i) + Re(conj(z) * w_j)
        acc += zr * wir - zi * wii + zr * wjr + zi * wji;
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

static int exp_blocks(uint64_t work) {
    uint64_t want = (work + kExpThreads * 4 - 1) / (kExpThreads * 4);
    if (want < 1) want = 1;
    if (want > kExpBlocksMax) want = kExpBlocksMax;
    return (int)want;
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
    int nb_diag = exp_blocks(st->dim);
    int nb_flip = exp_blocks(st->dim >> 1);
    int nb = std::max(nb_diag, nb_flip);
    size_t part_count = (size_t)ng * nb;
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
    std::memcpy(h, ps.terms.data(), term_bytes);
    LQ_HIP_CHECK(hipMemcpyAsync(d_terms, h, term_bytes, hipMemcpyHostToDevice, stream()));
    int written = 0;
    for (int g = 0; g < ng; ++g) {
        const PauliGroup& grp = ps.groups[g];
        if (grp.x == 0) {
            hipLaunchKernelGGL(k_expect_diag<V>, dim3(nb_diag), dim3(kExpThreads), 0, stream(),
                               (const V*)st->d_amps, st->dim, (const PauliTermDev*)(d_terms + grp.first), grp.count,
                               d_part + written);
            written += nb_diag;
        } else {
            int hbit = 63 - __builtin_clzll(grp.x);
            hipLaunchKernelGGL(k_expect_flip<V>, dim3(nb_flip), dim3(kExpThreads), 0, stream(),
                               (const V*)st->d_amps, st->dim >> 1, grp.x, hbit,
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
