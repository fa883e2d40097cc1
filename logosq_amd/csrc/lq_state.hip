// State storage, library set-up and the f64 reductions behind State::normalize
// and the inner products (reference: src/states/mod.rs:38-392).
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <atomic>
#include <mutex>
#include <vector>

#include "lq_common.h"

namespace lq {

// ---- errors -------------------------------------------------------------------
static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

void clear_error() { g_err[0] = 0; }

std::recursive_mutex& api_mutex() {
    static std::recursive_mutex m;
    return m;
}

// ---- library state ----------------------------------------------------------------
static std::mutex g_mu;
static bool g_inited = false;
// Written once under g_mu; read lock-free by stream() on every launch.
static std::atomic<hipStream_t> g_stream{nullptr};
static void* g_scratch = nullptr;
static size_t g_scratch_bytes = 0;
static void* g_scratch2 = nullptr;
static size_t g_scratch2_bytes = 0;
static double* g_host = nullptr;
static size_t g_host_count = 0;

int ensure_init() {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_inited) return LQ_OK;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) {
        set_error("no HIP device available (%s)", hipGetErrorString(e));
        return LQ_ERR_DEVICE;
    }
    // Default-priority, non-blocking stream: all library work is ordered on it.
    hipStream_t s = nullptr;
    LQ_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    g_stream.store(s, std::memory_order_release);
    g_inited = true;
    return LQ_OK;
}

// The library stream, created at the first use by any entry point (work put
// on the legacy null stream instead would run, but could not be captured
// into a graph).  nullptr only when no device can be initialised.
hipStream_t stream() {
    hipStream_t s = g_stream.load(std::memory_order_acquire);
    if (s) return s;
    (void)ensure_init();
    return g_stream.load(std::memory_order_acquire);
}

static void* grow(void*& buf, size_t& have, size_t bytes) {
    if (bytes <= have) return buf;
    if (buf) {
        if (hipStreamSynchronize(stream()) != hipSuccess) return nullptr;
        (void)hipFree(buf);
        buf = nullptr;
        have = 0;
    }
    size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
    if (hipMalloc(&buf, want) != hipSuccess) {
        buf = nullptr;
        return nullptr;
    }
    have = want;
    return buf;
}

void* scratch(size_t bytes) { return grow(g_scratch, g_scratch_bytes, bytes); }
void* scratch2(size_t bytes) { return grow(g_scratch2, g_scratch2_bytes, bytes); }

double* host_staging(size_t count) {
    if (count <= g_host_count) return g_host;
    if (g_host) (void)hipHostFree(g_host);
    g_host = nullptr;
    g_host_count = 0;
    size_t want = count < 4096 ? 4096 : count;
    if (hipHostMalloc((void**)&g_host, want * sizeof(double), hipHostMallocDefault) != hipSuccess) return nullptr;
    g_host_count = want;
    return g_host;
}

int check_state(const lq_state* st) {
    if (!st || !st->d_amps) {
        set_error("null or destroyed state handle");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    return LQ_OK;
}

int check_qubit(const lq_state* st, int q) {
    if (q < 0 || q >= st->n) {
        // LogosQError::InvalidQubitIndex text (src/error.rs:11-12)
        set_error("Invalid qubit index: %d (must be < %d)", q, st->n);
        return LQ_ERR_INVALID_QUBIT_INDEX;
    }
    return LQ_OK;
}

// ---- kernels --------------------------------------------------------------------------
constexpr int kRedBlocks = 2048;
constexpr int kRedThreads = 256;

template <int NT>
__device__ __forceinline__ double block_sum(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    return sh[0];
}

// Partial sums of |a|^2 in f64: fixed grid, fixed per-thread order, fixed tree:
// the result is bit-reproducible run to run.
template <typename V>
__global__ void __launch_bounds__(kRedThreads) k_sumsq(const V* __restrict__ a, uint64_t dim,
                                                       double* __restrict__ partial) {
    __shared__ double sh[kRedThreads];
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double s0 = 0.0, s1 = 0.0;
    for (; i + stride < dim; i += 2 * stride) {
        V x = a[i], y = a[i + stride];
        s0 += (double)x.x * (double)x.x + (double)x.y * (double)x.y;
        s1 += (double)y.x * (double)y.x + (double)y.y * (double)y.y;
    }
    if (i < dim) {
        V x = a[i];
        s0 += (double)x.x * (double)x.x + (double)x.y * (double)x.y;
    }
    double t = block_sum<kRedThreads>(s0 + s1, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// Sum `count` partials in a fixed order; optionally also write sqrt(sum).
__global__ void __launch_bounds__(1024) k_final_sum(const double* __restrict__ partial, int count,
                                                   double* __restrict__ out, int also_sqrt) {
    __shared__ double sh[1024];
    double s = 0.0;
    for (int i = threadIdx.x; i < count; i += blockDim.x) s += partial[i];
    double t = block_sum<1024>(s, sh);
    if (threadIdx.x == 0) {
        out[0] = t;
        if (also_sqrt) out[1] = sqrt(t);
    }
}

// a /= norm when norm > 1e-10 (src/states/mod.rs:101-103).  The division is
// done in f64 for both amplitude types.
// Streaming kernels take one element per thread and as many blocks as the
// state needs: blocks are dispatched in address order, so the waves in flight
// at any moment cover one contiguous window of memory (few pages to
// translate, whole DRAM rows).  On MI355X a 2 x 16 GiB copy runs at 6.6 TB/s
// this way, 4.6-5.3 TB/s as a grid-stride loop and 4.9 TB/s through
// hipMemcpyAsync.
template <typename V>
__global__ void __launch_bounds__(256) k_scale(V* __restrict__ a, uint64_t dim, const double* __restrict__ norm_ptr) {
    const double norm = norm_ptr[1];
    if (!(norm > 1e-10)) return;
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < dim) {
        V x = a[i];
        x.x = (decltype(x.x))((double)x.x / norm);
        x.y = (decltype(x.y))((double)x.y / norm);
        a[i] = x;
    }
}

template <typename V>
__global__ void __launch_bounds__(256) k_fill(V* __restrict__ a, uint64_t dim, double value, uint64_t one_index) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < dim) {
        V x;
        x.x = (decltype(x.x))(i == one_index ? 1.0 : value);
        x.y = 0;
        a[i] = x;
    }
}

typedef double copy_unit __attribute__((ext_vector_type(2)));

// Bit copy in 16-byte units (two complex64 or one complex128 amplitude).
__global__ void __launch_bounds__(256) k_copy16(const copy_unit* __restrict__ src, copy_unit* __restrict__ dst,
                                                uint64_t units) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < units) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

// Blocks of 256 threads covering `count` elements, one per thread.
static unsigned blocks_one_per_thread(uint64_t count) { return (unsigned)((count + 255) / 256); }

template <typename V>
__global__ void __launch_bounds__(kRedThreads) k_inner(const V* __restrict__ a, const V* __restrict__ b,
                                                       uint64_t dim, double* __restrict__ partial) {
    __shared__ double sh[kRedThreads];
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    double re = 0.0, im = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
        V x = a[i], y = b[i];
        // conj(x) * y
        re += (double)x.x * (double)y.x + (double)x.y * (double)y.y;
        im += (double)x.x * (double)y.y - (double)x.y * (double)y.x;
    }
    double tr = block_sum<kRedThreads>(re, sh);
    __syncthreads();
    double ti = block_sum<kRedThreads>(im, sh);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = tr;
        partial[2 * blockIdx.x + 1] = ti;
    }
}

__global__ void __launch_bounds__(1024) k_final_sum2(const double* __restrict__ partial, int count,
                                                    double* __restrict__ out) {
    __shared__ double sh[1024];
    double re = 0.0, im = 0.0;
    for (int i = threadIdx.x; i < count; i += blockDim.x) {
        re += partial[2 * i];
        im += partial[2 * i + 1];
    }
    double tr = block_sum<1024>(re, sh);
    __syncthreads();
    double ti = block_sum<1024>(im, sh);
    if (threadIdx.x == 0) {
        out[0] = tr;
        out[1] = ti;
    }
}

template <typename V>
__global__ void k_probs(const V* __restrict__ a, uint64_t dim, double* __restrict__ out) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
        V x = a[i];
        out[i] = (double)x.x * (double)x.x + (double)x.y * (double)x.y;
    }
}

__global__ void k_f64_to_f32(const double* __restrict__ in, float* __restrict__ out, uint64_t count) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) out[i] = (float)in[i];
}

static unsigned grid_for(uint64_t work, unsigned threads, unsigned cap = 65536u) {
    uint64_t g = (work + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

static int red_blocks(uint64_t dim) {
    uint64_t want = (dim + kRedThreads * 8 - 1) / (kRedThreads * 8);
    if (want < 1) want = 1;
    if (want > kRedBlocks) want = kRedBlocks;
    return (int)want;
}

// Leaves sum |a|^2 in d_out[0] and its sqrt in d_out[1].
int reduce_norm_sqr(const lq_state* st, double* d_out) {
    int nb = red_blocks(st->dim);
    double* partial = (double*)scratch((size_t)nb * sizeof(double));
    if (!partial) {
        set_error("scratch allocation failed");
        return LQ_ERR_OUT_OF_MEMORY;
    }
    if (st->dtype == LQ_C128)
        hipLaunchKernelGGL(k_sumsq<double2>, dim3(nb), dim3(kRedThreads), 0, stream(), (const double2*)st->d_amps,
                           st->dim, partial);
    else
        hipLaunchKernelGGL(k_sumsq<float2>, dim3(nb), dim3(kRedThreads), 0, stream(), (const float2*)st->d_amps,
                           st->dim, partial);
    LQ_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(1024), 0, stream(), (const double*)partial, nb, d_out, 1);
    LQ_LAUNCH_CHECK();
    return LQ_OK;
}

static double* g_norm_dev = nullptr;

static double* norm_slot() {
    if (!g_norm_dev) {
        if (hipMalloc((void**)&g_norm_dev, 4 * sizeof(double)) != hipSuccess) g_norm_dev = nullptr;
    }
    return g_norm_dev;
}

int normalize_device(lq_state* st, double* norm_out) {
    double* d = norm_slot();
    if (!d) {
        set_error("allocation of the norm slot failed");
        return LQ_ERR_OUT_OF_MEMORY;
    }
    int rc = reduce_norm_sqr(st, d);
    if (rc) return rc;
    const unsigned g = blocks_one_per_thread(st->dim);
    if (st->dtype == LQ_C128)
        hipLaunchKernelGGL(k_scale<double2>, dim3(g), dim3(256), 0, stream(), (double2*)st->d_amps, st->dim,
                           (const double*)d);
    else
        hipLaunchKernelGGL(k_scale<float2>, dim3(g), dim3(256), 0, stream(), (float2*)st->d_amps, st->dim,
                           (const double*)d);
    LQ_LAUNCH_CHECK();
    if (norm_out) {
        double* h = host_staging(4);
        if (!h) return LQ_ERR_OUT_OF_MEMORY;
        LQ_HIP_CHECK(hipMemcpyAsync(h, d, 2 * sizeof(double), hipMemcpyDeviceToHost, stream()));
        LQ_HIP_CHECK(hipStreamSynchronize(stream()));
        *norm_out = h[1];
    }
    return LQ_OK;
}

static int fill_state(lq_state* st, double value, uint64_t one_index) {
    const unsigned g = blocks_one_per_thread(st->dim);
    if (st->dtype == LQ_C128)
        hipLaunchKernelGGL(k_fill<double2>, dim3(g), dim3(256), 0, stream(), (double2*)st->d_amps, st->dim, value,
                           one_index);
    else
        hipLaunchKernelGGL(k_fill<float2>, dim3(g), dim3(256), 0, stream(), (float2*)st->d_amps, st->dim, value,
                           one_index);
    LQ_LAUNCH_CHECK();
    return LQ_OK;
}

}  // namespace lq

using namespace lq;

extern "C" {

const char* lq_last_error(void) { return g_err; }

int lq_abi_version(void) { return LQ_ABI_VERSION; }

int lq_device_info(char* name, size_t name_len, int* cu_count, size_t* lds_bytes, size_t* global_bytes) {
    LQ_API_GUARD;
    int rc = ensure_init();
    if (rc) return rc;
    int dev = 0;
    LQ_HIP_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t p;
    LQ_HIP_CHECK(hipGetDeviceProperties(&p, dev));
    if (name && name_len) {
        snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
    }
    if (cu_count) *cu_count = p.multiProcessorCount;
    if (lds_bytes) *lds_bytes = p.sharedMemPerBlock;
    if (global_bytes) *global_bytes = p.totalGlobalMem;
    return LQ_OK;
}

int lq_synchronize(void) {
    LQ_API_GUARD;
    int rc = ensure_init();
    if (rc) return rc;
    LQ_HIP_CHECK(hipStreamSynchronize(stream()));
    return LQ_OK;
}

static hipEvent_t g_events[64];
static bool g_event_made[64];

int lq_event_record(int slot) {
    LQ_API_GUARD;
    int rc = ensure_init();
    if (rc) return rc;
    if (slot < 0 || slot >= 64) {
        set_error("event slot %d outside [0, 64)", slot);
        return LQ_ERR_INVALID_ARGUMENT;
    }
    if (!g_event_made[slot]) {
        LQ_HIP_CHECK(hipEventCreate(&g_events[slot]));
        g_event_made[slot] = true;
    }
    LQ_HIP_CHECK(hipEventRecord(g_events[slot], stream()));
    return LQ_OK;
}

int lq_event_elapsed_ms(int a, int b, float* ms) {
    LQ_API_GUARD;
    if (a < 0 || a >= 64 || b < 0 || b >= 64 || !g_event_made[a] || !g_event_made[b] || !ms) {
        set_error("event slots not recorded");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    LQ_HIP_CHECK(hipEventSynchronize(g_events[b]));
    LQ_HIP_CHECK(hipEventElapsedTime(ms, g_events[a], g_events[b]));
    return LQ_OK;
}

int lq_state_create(int num_qubits, int dtype, lq_state** out) {
    LQ_API_GUARD;
    clear_error();
    if (!out) {
        set_error("out pointer is null");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    if (num_qubits < 1 || num_qubits > kMaxQubits) {
        set_error("num_qubits %d outside [1, %d]", num_qubits, kMaxQubits);
        return LQ_ERR_INVALID_ARGUMENT;
    }
    if (dtype != LQ_C128 && dtype != LQ_C64) {
        set_error("unknown dtype %d", dtype);
        return LQ_ERR_INVALID_ARGUMENT;
    }
    int rc = ensure_init();
    if (rc) return rc;
    lq_state* st = new lq_state();
    st->n = num_qubits;
    st->dtype = dtype;
    st->dim = 1ull << num_qubits;
    st->amp_bytes = dtype == LQ_C128 ? 16 : 8;
    hipError_t e = hipMalloc(&st->d_amps, st->dim * st->amp_bytes);
    if (e != hipSuccess) {
        set_error("hipMalloc of %llu bytes failed: %s", (unsigned long long)(st->dim * st->amp_bytes),
                  hipGetErrorString(e));
        (void)hipGetLastError();
        delete st;
        return LQ_ERR_OUT_OF_MEMORY;
    }
    rc = fill_state(st, 0.0, 0);
    if (rc) {
        (void)hipFree(st->d_amps);
        delete st;
        return rc;
    }
    *out = st;
    return LQ_OK;
}

int lq_state_destroy(lq_state* st) {
    LQ_API_GUARD;
    if (!st) return LQ_OK;
    if (st->d_amps) {
        (void)hipStreamSynchronize(stream());
        (void)hipFree(st->d_amps);
    }
    st->d_amps = nullptr;
    delete st;
    return LQ_OK;
}

int lq_state_num_qubits(const lq_state* st, int* num_qubits) {
    LQ_API_GUARD;
    int rc = check_state(st);
    if (rc) return rc;
    *num_qubits = st->n;
    return LQ_OK;
}

int lq_state_dtype(const lq_state* st, int* dtype) {
    LQ_API_GUARD;
    int rc = check_state(st);
    if (rc) return rc;
    *dtype = st->dtype;
    return LQ_OK;
}

int lq_state_set_basis(lq_state* st, uint64_t index) {
    LQ_API_GUARD;
    clear_error();
    int rc = check_state(st);
    if (rc) return rc;
    if (index >= st->dim) {
        set_error("basis index %llu out of range for %d qubits", (unsigned long long)index, st->n);
        return LQ_ERR_INVALID_ARGUMENT;
    }
    return fill_state(st, 0.0, index);
}

int lq_state_set_plus(lq_state* st) {
    LQ_API_GUARD;
    clear_error();
    int rc = check_state(st);
    if (rc) return rc;
    double a = 1.0 / std::sqrt((double)st->dim);  // src/states/mod.rs:154
    return fill_state(st, a, ~0ull);
}

int lq_state_upload(lq_state* st, const double* amps, uint64_t len, int normalize) {
    LQ_API_GUARD;
    clear_error();
    int rc = check_state(st);
    if (rc) return rc;
    if (!amps) {
        set_error("amps is null");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    if (len != st->dim) {
        set_error("Invalid state dimension: %llu (must be a power of 2)", (unsigned long long)len);
        return LQ_ERR_INVALID_STATE_DIMENSION;
    }
    if (st->dtype == LQ_C128) {
        LQ_HIP_CHECK(hipMemcpyAsync(st->d_amps, amps, len * 16, hipMemcpyHostToDevice, stream()));
    } else {
        // stage f64 on the device, convert to f32 there
        uint64_t count = 2 * len;
        double* tmp = (double*)scratch2(count * sizeof(double));
        if (!tmp) {
            set_error("staging allocation failed");
            return LQ_ERR_OUT_OF_MEMORY;
        }
        LQ_HIP_CHECK(hipMemcpyAsync(tmp, amps, count * sizeof(double), hipMemcpyHostToDevice, stream()));
        unsigned g = grid_for(count, 256, 16384);
        hipLaunchKernelGGL(k_f64_to_f32, dim3(g), dim3(256), 0, stream(), (const double*)tmp, (float*)st->d_amps,
                           count);
        LQ_LAUNCH_CHECK();
    }
    LQ_HIP_CHECK(hipStreamSynchronize(stream()));
    if (normalize) return normalize_device(st, nullptr);
    return LQ_OK;
}

int lq_state_download(const lq_state* st, double* amps, uint64_t len) {
    LQ_API_GUARD;
    clear_error();
    int rc = check_state(st);
    if (rc) return rc;
    if (!amps || len != st->dim) {
        set_error("download buffer must hold exactly 2^n = %llu amplitudes", (unsigned long long)st->dim);
        return LQ_ERR_DIMENSION_MISMATCH;
    }
    if (st->dtype == LQ_C128) {
        LQ_HIP_CHECK(hipMemcpyAsync(amps, st->d_amps, len * 16, hipMemcpyDeviceToHost, stream()));
        LQ_HIP_CHECK(hipStreamSynchronize(stream()));
    } else {
        std::vector<float> tmp(2 * len);
        LQ_HIP_CHECK(hipMemcpyAsync(tmp.data(), st->d_amps, len * 8, hipMemcpyDeviceToHost, stream()));
        LQ_HIP_CHECK(hipStreamSynchronize(stream()));
        for (uint64_t i = 0; i < 2 * len; ++i) amps[i] = (double)tmp[i];
    }
    return LQ_OK;
}

int lq_state_copy(lq_state* dst, const lq_state* src) {
    LQ_API_GUARD;
    clear_error();
    int rc = check_state(dst);
    if (rc) return rc;
    rc = check_state(src);
    if (rc) return rc;
    if (dst->n != src->n || dst->dtype != src->dtype) {
        set_error("Dimension mismatch: expected %llu, got %llu", (unsigned long long)dst->dim,
                  (unsigned long long)src->dim);
        return LQ_ERR_DIMENSION_MISMATCH;
    }
    const uint64_t bytes = src->dim * src->amp_bytes;
    if (bytes % 16 == 0) {
        hipLaunchKernelGGL(k_copy16, dim3(blocks_one_per_thread(bytes / 16)), dim3(256), 0, stream(),
                           (const copy_unit*)src->d_amps, (copy_unit*)dst->d_amps, bytes / 16);
        LQ_LAUNCH_CHECK();
    } else {
        LQ_HIP_CHECK(hipMemcpyAsync(dst->d_amps, src->d_amps, bytes, hipMemcpyDeviceToDevice, stream()));
    }
    return LQ_OK;
}

int lq_state_normalize(lq_state* st, double* norm_out) {
    LQ_API_GUARD;
    clear_error();
    int rc = check_state(st);
    if (rc) return rc;
    return normalize_device(st, norm_out);
}

int lq_state_norm_sqr(const lq_state* st, double* out) {
    LQ_API_GUARD;
    clear_error();
    int rc = check_state(st);
    if (rc) return rc;
    double* d = norm_slot();
    if (!d) return LQ_ERR_OUT_OF_MEMORY;
    rc = reduce_norm_sqr(st, d);
    if (rc) return rc;
    double* h = host_staging(4);
    if (!h) return LQ_ERR_OUT_OF_MEMORY;
    LQ_HIP_CHECK(hipMemcpyAsync(h, d, sizeof(double), hipMemcpyDeviceToHost, stream()));
    LQ_HIP_CHECK(hipStreamSynchronize(stream()));
    *out = h[0];
    return LQ_OK;
}

int lq_state_inner_product(const lq_state* a, const lq_state* b, double out[2]) {
    LQ_API_GUARD;
    clear_error();
    int rc = check_state(a);
    if (rc) return rc;
    rc = check_state(b);
    if (rc) return rc;
    if (a->dim != b->dim || a->dtype != b->dtype) {
        // DimensionMismatch (src/states/mod.rs:353-358)
        set_error("Dimension mismatch: expected %llu, got %llu", (unsigned long long)a->dim,
                  (unsigned long long)b->dim);
        return LQ_ERR_DIMENSION_MISMATCH;
    }
    int nb = red_blocks(a->dim);
    double* partial = (double*)scratch((size_t)nb * 2 * sizeof(double));
    double* d = norm_slot();
    if (!partial || !d) return LQ_ERR_OUT_OF_MEMORY;
    if (a->dtype == LQ_C128)
        hipLaunchKernelGGL(k_inner<double2>, dim3(nb), dim3(kRedThreads), 0, stream(), (const double2*)a->d_amps,
                           (const double2*)b->d_amps, a->dim, partial);
    else
        hipLaunchKernelGGL(k_inner<float2>, dim3(nb), dim3(kRedThreads), 0, stream(), (const float2*)a->d_amps,
                           (const float2*)b->d_amps, a->dim, partial);
    LQ_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_final_sum2, dim3(1), dim3(1024), 0, stream(), (const double*)partial, nb, d);
    LQ_LAUNCH_CHECK();
    double* h = host_staging(4);
    if (!h) return LQ_ERR_OUT_OF_MEMORY;
    LQ_HIP_CHECK(hipMemcpyAsync(h, d, 2 * sizeof(double), hipMemcpyDeviceToHost, stream()));
    LQ_HIP_CHECK(hipStreamSynchronize(stream()));
    out[0] = h[0];
    out[1] = h[1];
    return LQ_OK;
}

int lq_state_probabilities(const lq_state* st, double* out, uint64_t len) {
    LQ_API_GUARD;
    clear_error();
    int rc = check_state(st);
    if (rc) return rc;
    if (!out || len != st->dim) {
        set_error("probabilities buffer must hold exactly 2^n doubles");
        return LQ_ERR_DIMENSION_MISMATCH;
    }
    double* d = (double*)scratch2(st->dim * sizeof(double));
    if (!d) return LQ_ERR_OUT_OF_MEMORY;
    unsigned g = grid_for(st->dim, 256, 16384);
    if (st->dtype == LQ_C128)
        hipLaunchKernelGGL(k_probs<double2>, dim3(g), dim3(256), 0, stream(), (const double2*)st->d_amps, st->dim, d);
    else
        hipLaunchKernelGGL(k_probs<float2>, dim3(g), dim3(256), 0, stream(), (const float2*)st->d_amps, st->dim, d);
    LQ_LAUNCH_CHECK();
    LQ_HIP_CHECK(hipMemcpyAsync(out, d, st->dim * sizeof(double), hipMemcpyDeviceToHost, stream()));
    LQ_HIP_CHECK(hipStreamSynchronize(stream()));
    return LQ_OK;
}

}  // extern "C"
