// Measurement on the device: State::measure / measure_qubit /
// measure_shots_parallel and Circuit::execute_and_measure
// (src/states/mod.rs:180-262, src/circuits/mod.rs:215-247).
//
// The reference draws from thread_rng (unseeded); here the caller passes the
// uniforms (lq_measure_qubit, lq_bit_probabilities) or a seed for a
// SplitMix64 stream (lq_sample_shots), so results are reproducible.
//
// Shot sampling follows WeightedIndex: with cumulative weights c_i and total
// W, a uniform u gives x = u * W and the index is the first i with c_i > x.
// The state is cut into blocks of 4096 amplitudes; the host prefix-sums the
// block weights, sorts the shots' x values and hands each block the run of
// shots that fall into it; a block scans its own weights in LDS once and
// answers its shots by binary search.  Traffic: one read of the state for the
// block weights and one more for the scans, whatever the number of shots.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "lq_common.h"

namespace lq {

constexpr int kMThreads = 256;
constexpr int kSampleBlockBits = 12;  // 4096 amplitudes per sampling block
constexpr uint32_t kSampleBlock = 1u << kSampleBlockBits;

// p1[b] partials: per block, sum of |a_i|^2 over i with bit b set, b < n.
constexpr int kBPThreads = 128;
template <typename V>
__global__ void __launch_bounds__(kBPThreads) k_bit_probs(const V* __restrict__ a, uint64_t dim, int n,
                                                          double* __restrict__ partial) {
    __shared__ double sh[kMaxQubits][kBPThreads];
    double acc[kMaxQubits];
#pragma unroll
    for (int b = 0; b < kMaxQubits; ++b) acc[b] = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * kBPThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kBPThreads + threadIdx.x; i < dim; i += stride) {
        const V x = a[i];
        const double p = (double)x.x * (double)x.x + (double)x.y * (double)x.y;
#pragma unroll
        for (int b = 0; b < kMaxQubits; ++b) acc[b] += ((i >> b) & 1ull) ? p : 0.0;
    }
#pragma unroll
    for (int b = 0; b < kMaxQubits; ++b) sh[b][threadIdx.x] = acc[b];
    __syncthreads();
    if ((int)threadIdx.x < n) {
        double t = 0.0;
        for (int k = 0; k < kBPThreads; ++k) t += sh[threadIdx.x][k];
        partial[(uint64_t)blockIdx.x * n + threadIdx.x] = t;
    }
}

// out[b] = sum over blocks of partial[blk * n + b], fixed order.
__global__ void k_bit_probs_final(const double* __restrict__ partial, int nblocks, int n, double* __restrict__ out) {
    const int b = threadIdx.x;
    if (b >= n) return;
    double s = 0.0;
    for (int k = 0; k < nblocks; ++k) s += partial[(uint64_t)k * n + b];
    out[b] = s;
}

// Keep the amplitudes whose bit `bit` equals `value`, zero the others.
template <typename V>
__global__ void k_project(V* __restrict__ a, uint64_t dim, int bit, int value) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride)
        if ((int)((i >> bit) & 1ull) != value) {
            V z;
            z.x = 0;
            z.y = 0;
            a[i] = z;
        }
}

// Weight of each sampling block (tree sum of |a|^2 over its amplitudes).
template <typename V>
__global__ void __launch_bounds__(kMThreads) k_block_weights(const V* __restrict__ a, double* __restrict__ w) {
    __shared__ double sh[kMThreads];
    const V* blk = a + ((uint64_t)blockIdx.x << kSampleBlockBits);
    double s = 0.0;
    for (uint32_t i = threadIdx.x; i < kSampleBlock; i += kMThreads) {
        const V x = blk[i];
        s += (double)x.x * (double)x.x + (double)x.y * (double)x.y;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int k = kMThreads / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) w[blockIdx.x] = sh[0];
}

// Block b answers shots [range[b], range[b+1]) of the sorted x values: the
// inclusive scan of its weights (offset by the block's prefix) in LDS, then
// for each shot the first element whose cumulative weight exceeds x.
template <typename V>
__global__ void __launch_bounds__(kMThreads) k_sample_blocks(const V* __restrict__ a, uint32_t block_count,
                                                             const double* __restrict__ prefix,
                                                             const int64_t* __restrict__ range,
                                                             const double* __restrict__ xs,
                                                             const int64_t* __restrict__ shot_of,
                                                             uint64_t* __restrict__ out) {
    __shared__ double cdf[kSampleBlock];
    __shared__ double part[kMThreads];
    const uint32_t b = blockIdx.x;
    const int64_t s0 = range[b], s1 = range[b + 1];
    if (s0 == s1) return;  // uniform per block: no barrier is skipped by part of it
    const V* blk = a + ((uint64_t)b << kSampleBlockBits);
    constexpr int PER = kSampleBlock / kMThreads;  // 16 consecutive elements per thread
    const uint32_t base = threadIdx.x * PER;
    double run = 0.0;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const V x = blk[base + j];
        run += (double)x.x * (double)x.x + (double)x.y * (double)x.y;
        cdf[base + j] = run;
    }
    part[threadIdx.x] = run;
    __syncthreads();
    // exclusive scan of the per-thread totals (Hillis-Steele on 256 values)
    for (int off = 1; off < kMThreads; off <<= 1) {
        const double v = threadIdx.x >= (uint32_t)off ? part[threadIdx.x - off] : 0.0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    const double before = prefix[b] + (threadIdx.x ? part[threadIdx.x - 1] : 0.0);
#pragma unroll
    for (int j = 0; j < PER; ++j) cdf[base + j] += before;
    __syncthreads();
    for (int64_t s = s0 + threadIdx.x; s < s1; s += kMThreads) {
        const double x = xs[s];
        // first element with cdf > x (the last element if rounding leaves none)
        uint32_t lo = 0, hi = kSampleBlock - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (cdf[mid] > x) hi = mid;
            else lo = mid + 1;
        }
        out[shot_of[s]] = ((uint64_t)b << kSampleBlockBits) | lo;
    }
    (void)block_count;
}

static uint64_t splitmix(uint64_t& s) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static int bit_probs(const lq_state* st, double* p1_host) {
    const int n = st->n;
    const int nblocks = (int)std::min<uint64_t>(std::max<uint64_t>(st->dim / (kBPThreads * 64), 1), 2048);
    double* part = (double*)scratch(((size_t)nblocks * n + n) * sizeof(double));
    if (!part) {
        set_error("scratch allocation failed");
        return LQ_ERR_OUT_OF_MEMORY;
    }
    double* out = part + (size_t)nblocks * n;
    if (st->dtype == LQ_C128)
        hipLaunchKernelGGL(k_bit_probs<double2>, dim3(nblocks), dim3(kBPThreads), 0, stream(),
                           (const double2*)st->d_amps, st->dim, n, part);
    else
        hipLaunchKernelGGL(k_bit_probs<float2>, dim3(nblocks), dim3(kBPThreads), 0, stream(),
                           (const float2*)st->d_amps, st->dim, n, part);
    hipLaunchKernelGGL(k_bit_probs_final, dim3(1), dim3(64), 0, stream(), (const double*)part, nblocks, n, out);
    LQ_LAUNCH_CHECK();
    double* h = host_staging(n);
    if (!h) return LQ_ERR_OUT_OF_MEMORY;
    LQ_HIP_CHECK(hipMemcpyAsync(h, out, n * sizeof(double), hipMemcpyDeviceToHost, stream()));
    LQ_HIP_CHECK(hipStreamSynchronize(stream()));
    std::memcpy(p1_host, h, n * sizeof(double));
    return LQ_OK;
}

}  // namespace lq

using namespace lq;

extern "C" {

int lq_bit_probabilities(const lq_state* st, double* p1) {
    LQ_API_GUARD;
    clear_error();
    int rc = check_state(st);
    if (rc) return rc;
    if (!p1) {
        set_error("output is null");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    return bit_probs(st, p1);
}

int lq_measure_qubit(lq_state* st, int qubit_idx, double u, int* result) {
    LQ_API_GUARD;
    clear_error();
    int rc = check_state(st);
    if (rc) return rc;
    if (qubit_idx < 0 || qubit_idx >= st->n) {
        set_error("Invalid qubit index: %d (must be < %d)", qubit_idx, st->n);
        return LQ_ERR_INVALID_QUBIT_INDEX;
    }
    if (!result) {
        set_error("result is null");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    std::vector<double> p1(st->n);
    rc = bit_probs(st, p1.data());
    if (rc) return rc;
    // mask = 1 << qubit_idx (src/states/mod.rs:205): bit qubit_idx of the index
    const int bit = qubit_idx;
    const int r = u < p1[bit] ? 1 : 0;
    const unsigned g = (unsigned)std::min<uint64_t>((st->dim + 255) / 256, 16384);
    if (st->dtype == LQ_C128)
        hipLaunchKernelGGL(k_project<double2>, dim3(g), dim3(256), 0, stream(), (double2*)st->d_amps, st->dim, bit, r);
    else
        hipLaunchKernelGGL(k_project<float2>, dim3(g), dim3(256), 0, stream(), (float2*)st->d_amps, st->dim, bit, r);
    LQ_LAUNCH_CHECK();
    rc = normalize_device(st, nullptr);
    if (rc) return rc;
    *result = r;
    return LQ_OK;
}

int lq_sample_shots(const lq_state* st, uint64_t seed, int64_t shots, uint64_t* out) {
    LQ_API_GUARD;
    clear_error();
    int rc = check_state(st);
    if (rc) return rc;
    if (shots < 0 || (shots > 0 && !out)) {
        set_error("invalid shot buffer");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    if (shots == 0) return LQ_OK;
    if (st->n < kSampleBlockBits) {
        // small states: the whole distribution on the host, same rule
        std::vector<double> amps(2 * st->dim);
        rc = lq_state_download(st, amps.data(), st->dim);
        if (rc) return rc;
        std::vector<double> c(st->dim);
        double run = 0.0;
        for (uint64_t i = 0; i < st->dim; ++i) {
            run += amps[2 * i] * amps[2 * i] + amps[2 * i + 1] * amps[2 * i + 1];
            c[i] = run;
        }
        uint64_t s = seed;
        for (int64_t k = 0; k < shots; ++k) {
            const double x = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0) * run;
            uint64_t i = (uint64_t)(std::upper_bound(c.begin(), c.end(), x) - c.begin());
            out[k] = std::min<uint64_t>(i, st->dim - 1);
        }
        return LQ_OK;
    }
    const uint32_t nb = (uint32_t)(st->dim >> kSampleBlockBits);
    // block weights -> host prefix
    double* d_w = (double*)scratch2((size_t)nb * sizeof(double));
    if (!d_w) return LQ_ERR_OUT_OF_MEMORY;
    if (st->dtype == LQ_C128)
        hipLaunchKernelGGL(k_block_weights<double2>, dim3(nb), dim3(kMThreads), 0, stream(),
                           (const double2*)st->d_amps, d_w);
    else
        hipLaunchKernelGGL(k_block_weights<float2>, dim3(nb), dim3(kMThreads), 0, stream(),
                           (const float2*)st->d_amps, d_w);
    LQ_LAUNCH_CHECK();
    std::vector<double> w(nb);
    LQ_HIP_CHECK(hipMemcpyAsync(w.data(), d_w, nb * sizeof(double), hipMemcpyDeviceToHost, stream()));
    LQ_HIP_CHECK(hipStreamSynchronize(stream()));
    std::vector<double> prefix(nb + 1, 0.0);
    for (uint32_t b = 0; b < nb; ++b) prefix[b + 1] = prefix[b] + w[b];
    const double total = prefix[nb];
    // shots: x = u * total, sorted, with their original positions
    std::vector<double> xs(shots);
    uint64_t s = seed;
    for (int64_t k = 0; k < shots; ++k) xs[k] = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0) * total;
    std::vector<int64_t> order(shots);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int64_t i, int64_t j) { return xs[i] < xs[j]; });
    std::vector<double> xsorted(shots);
    for (int64_t k = 0; k < shots; ++k) xsorted[k] = xs[order[k]];
    // block b takes the shots with prefix[b] <= x < prefix[b+1] (the last
    // block with weight takes any x at or beyond the total)
    std::vector<int64_t> range(nb + 1, 0);
    {
        int64_t k = 0;
        uint32_t last = 0;
        for (uint32_t b = 0; b < nb; ++b)
            if (w[b] > 0.0) last = b;
        for (uint32_t b = 0; b < nb; ++b) {
            range[b] = k;
            if (b == last) k = shots;
            else
                while (k < shots && xsorted[k] < prefix[b + 1]) ++k;
        }
        range[nb] = shots;
    }
    // device buffers: prefix, range, xs, order, out
    const size_t bp = (size_t)(nb + 1) * 8, br = (size_t)(nb + 1) * 8, bx = (size_t)shots * 8, bo = (size_t)shots * 8;
    char* d = (char*)scratch(bp + br + bx + bo + bo + 1024);
    if (!d) return LQ_ERR_OUT_OF_MEMORY;
    double* d_prefix = (double*)d;
    int64_t* d_range = (int64_t*)(d + bp);
    double* d_xs = (double*)(d + bp + br);
    int64_t* d_order = (int64_t*)(d + bp + br + bx);
    uint64_t* d_out = (uint64_t*)(d + bp + br + bx + bo);
    LQ_HIP_CHECK(hipMemcpyAsync(d_prefix, prefix.data(), bp, hipMemcpyHostToDevice, stream()));
    LQ_HIP_CHECK(hipMemcpyAsync(d_range, range.data(), br, hipMemcpyHostToDevice, stream()));
    LQ_HIP_CHECK(hipMemcpyAsync(d_xs, xsorted.data(), bx, hipMemcpyHostToDevice, stream()));
    LQ_HIP_CHECK(hipMemcpyAsync(d_order, order.data(), bo, hipMemcpyHostToDevice, stream()));
    if (st->dtype == LQ_C128)
        hipLaunchKernelGGL(k_sample_blocks<double2>, dim3(nb), dim3(kMThreads), 0, stream(),
                           (const double2*)st->d_amps, nb, (const double*)d_prefix, (const int64_t*)d_range,
                           (const double*)d_xs, (const int64_t*)d_order, d_out);
    else
        hipLaunchKernelGGL(k_sample_blocks<float2>, dim3(nb), dim3(kMThreads), 0, stream(),
                           (const float2*)st->d_amps, nb, (const double*)d_prefix, (const int64_t*)d_range,
                           (const double*)d_xs, (const int64_t*)d_order, d_out);
    LQ_LAUNCH_CHECK();
    LQ_HIP_CHECK(hipMemcpyAsync(out, d_out, bo, hipMemcpyDeviceToHost, stream()));
    LQ_HIP_CHECK(hipStreamSynchronize(stream()));
    return LQ_OK;
}

}  // extern "C"
