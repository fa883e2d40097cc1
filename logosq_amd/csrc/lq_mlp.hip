// Warm-start network: a multilayer perceptron in -> H -> H -> out (GELU)
// trained and evaluated on the GPU.  No counterpart in the reference.
//
// Weights are kept as f32 masters (Adam state in f32) and every matrix
// product rounds its operands to bf16 (round to nearest even) and
// accumulates in f32 on the matrix cores: v_mfma_f32_16x16x16_bf16, one
// 16x16 output tile per wavefront and K consumed 16 at a time.  Layouts are
// row-major; a weight W of a layer with `in` inputs and `out` outputs is
// [out][in].  Parameter order: W1, b1, W2, b2, W3, b3.
#include <cmath>
#include <cstring>
#include <vector>

#include "lq_common.h"

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct lq_mlp {
    int in = 0, hid = 0, out = 0;
    int64_t nparams = 0;
    float* w = nullptr;     // device f32 master parameters
    float* g = nullptr;     // device gradients
    float* m = nullptr;     // Adam first moment
    float* v = nullptr;     // Adam second moment
    int64_t step = 0;
    // activations / work (capacity in rows)
    int cap = 0;
    float *x = nullptr, *t = nullptr, *z1 = nullptr, *a1 = nullptr, *z2 = nullptr, *a2 = nullptr, *y = nullptr;
    float *d3 = nullptr, *d2 = nullptr, *d1 = nullptr, *loss = nullptr;
};

namespace lq {

// f32 -> bf16 bits, round to nearest even (NaN kept quiet).
__device__ __forceinline__ short f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (short)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

__device__ __forceinline__ float bf2f(short b) { return __uint_as_float(((uint32_t)(unsigned short)b) << 16); }

// C[M x N] = opA(A)[M x K] * opB(B)[K x N] (+ bias[N]) with bf16 operands and
// f32 accumulation.  TA: A is stored K x M (use A^T); TB: B is stored N x K.
// Block = 4 wavefronts = a 32 x 32 block of C (each wave one 16 x 16 tile).
template <bool TA, bool TB>
__global__ void __launch_bounds__(256) k_gemm_bf16(const float* __restrict__ A, const float* __restrict__ B,
                                                   const float* __restrict__ bias, float* __restrict__ C, int M,
                                                   int N, int K, int lda, int ldb, int ldc, int accumulate) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int m0 = blockIdx.y * 32 + (wave >> 1) * 16;
    const int n0 = blockIdx.x * 32 + (wave & 1) * 16;
    const int r = lane & 15;        // A row / B column owned by this lane
    const int kq = (lane >> 4) * 4; // first of the 4 K indices of this lane
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 av, bv;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int k = k0 + kq + i;
            const int am = m0 + r, bn = n0 + r;
            float a = 0.f, b = 0.f;
            if (am < M && k < K) a = TA ? A[(size_t)k * lda + am] : A[(size_t)am * lda + k];
            if (bn < N && k < K) b = TB ? B[(size_t)bn * ldb + k] : B[(size_t)k * ldb + bn];
            av[i] = f2bf(a);
            bv[i] = f2bf(b);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
    // D layout: lane holds rows (lane >> 4) * 4 + i, column lane & 15
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int row = m0 + (lane >> 4) * 4 + i;
        const int col = n0 + (lane & 15);
        if (row < M && col < N) {
            float vv = acc[i] + (bias ? bias[col] : 0.f);
            if (accumulate) vv += C[(size_t)row * ldc + col];
            C[(size_t)row * ldc + col] = vv;
        }
    }
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_grad(float z) {
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
    return cdf + z * pdf;
}

__global__ void k_gelu(const float* __restrict__ z, float* __restrict__ a, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = gelu(z[i]);
}

// d = d_out * gelu'(z)
__global__ void k_gelu_back(const float* __restrict__ z, float* __restrict__ d, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] *= gelu_grad(z[i]);
}

// MSE: d3 = 2 (y - t) / (rows * cols); per-block partial losses.
__global__ void __launch_bounds__(256) k_mse(const float* __restrict__ y, const float* __restrict__ t,
                                             float* __restrict__ d3, int64_t n, float scale,
                                             float* __restrict__ partial) {
    __shared__ float sh[256];
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (i < n) {
        float e = y[i] - t[i];
        d3[i] = 2.0f * e * scale;
        l = e * e * scale;
    }
    sh[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void k_sum_f(const float* __restrict__ p, int n, float* __restrict__ out) {
    __shared__ float sh[256];
    float s = 0.f;
    for (int i = threadIdx.x; i < n; i += 256) s += p[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

// bias gradient: db[c] = sum_r d[r][c]
__global__ void k_colsum(const float* __restrict__ d, float* __restrict__ db, int rows, int cols) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.f;
    for (int r = 0; r < rows; ++r) s += d[(size_t)r * cols + c];
    db[c] = s;
}

// First layer (in <= 8): z1 = x W1^T + b1 in f32 (a 3-term dot product
// gains nothing from the matrix cores); weights rounded to bf16 like the
// other layers.
__global__ void k_layer1(const float* __restrict__ x, const float* __restrict__ W, const float* __restrict__ b,
                         float* __restrict__ z, int rows, int in, int hid) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * hid) return;
    int r = i / hid, c = i - r * hid;
    float s = 0.f;
    for (int k = 0; k < in; ++k) s += bf2f(f2bf(x[(size_t)r * in + k])) * bf2f(f2bf(W[(size_t)c * in + k]));
    z[i] = s + b[c];
}

// dW1[c][k] = sum_r d1[r][c] * x[r][k] (bf16-rounded operands, f32 sums)
__global__ void k_layer1_grad(const float* __restrict__ x, const float* __restrict__ d1, float* __restrict__ gW,
                              int rows, int in, int hid) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hid * in) return;
    int c = i / in, k = i - c * in;
    float s = 0.f;
    for (int r = 0; r < rows; ++r) s += bf2f(f2bf(d1[(size_t)r * hid + c])) * bf2f(f2bf(x[(size_t)r * in + k]));
    gW[i] = s;
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    float mi = b1 * m[i] + (1.f - b1) * gi;
    float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

static unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

template <bool TA, bool TB>
static void gemm(const float* A, const float* B, const float* bias, float* C, int M, int N, int K, int lda, int ldb,
                 int ldc, int acc) {
    dim3 grid((N + 31) / 32, (M + 31) / 32);
    hipLaunchKernelGGL((k_gemm_bf16<TA, TB>), grid, dim3(256), 0, stream(), A, B, bias, C, M, N, K, lda, ldb, ldc,
                       acc);
}

struct Off {
    int64_t W1, b1, W2, b2, W3, b3;
};

static Off offsets(const lq_mlp* n) {
    Off o;
    o.W1 = 0;
    o.b1 = o.W1 + (int64_t)n->hid * n->in;
    o.W2 = o.b1 + n->hid;
    o.b2 = o.W2 + (int64_t)n->hid * n->hid;
    o.W3 = o.b2 + n->hid;
    o.b3 = o.W3 + (int64_t)n->out * n->hid;
    return o;
}

static int ensure_cap(lq_mlp* n, int rows) {
    if (rows <= n->cap) return LQ_OK;
    float** bufs[] = {&n->x, &n->t, &n->z1, &n->a1, &n->z2, &n->a2, &n->y, &n->d3, &n->d2, &n->d1, &n->loss};
    for (float** b : bufs)
        if (*b) {
            (void)hipFree(*b);
            *b = nullptr;
        }
    size_t r = (size_t)rows;
    LQ_HIP_CHECK(hipMalloc((void**)&n->x, r * n->in * 4));
    LQ_HIP_CHECK(hipMalloc((void**)&n->t, r * n->out * 4));
    LQ_HIP_CHECK(hipMalloc((void**)&n->z1, r * n->hid * 4));
    LQ_HIP_CHECK(hipMalloc((void**)&n->a1, r * n->hid * 4));
    LQ_HIP_CHECK(hipMalloc((void**)&n->z2, r * n->hid * 4));
    LQ_HIP_CHECK(hipMalloc((void**)&n->a2, r * n->hid * 4));
    LQ_HIP_CHECK(hipMalloc((void**)&n->y, r * n->out * 4));
    LQ_HIP_CHECK(hipMalloc((void**)&n->d3, r * n->out * 4));
    LQ_HIP_CHECK(hipMalloc((void**)&n->d2, r * n->hid * 4));
    LQ_HIP_CHECK(hipMalloc((void**)&n->d1, r * n->hid * 4));
    LQ_HIP_CHECK(hipMalloc((void**)&n->loss, (r * n->out / 256 + 16) * 4));
    n->cap = rows;
    return LQ_OK;
}

// forward on device buffers n->x (rows) -> n->y
static int forward(lq_mlp* n, int rows) {
    Off o = offsets(n);
    const float* w = n->w;
    hipLaunchKernelGGL(k_layer1, dim3(nblk((int64_t)rows * n->hid)), dim3(256), 0, stream(), n->x, w + o.W1,
                       w + o.b1, n->z1, rows, n->in, n->hid);
    hipLaunchKernelGGL(k_gelu, dim3(nblk((int64_t)rows * n->hid)), dim3(256), 0, stream(), n->z1, n->a1,
                       (int64_t)rows * n->hid);
    // z2 = a1 W2^T + b2
    gemm<false, true>(n->a1, w + o.W2, w + o.b2, n->z2, rows, n->hid, n->hid, n->hid, n->hid, n->hid, 0);
    hipLaunchKernelGGL(k_gelu, dim3(nblk((int64_t)rows * n->hid)), dim3(256), 0, stream(), n->z2, n->a2,
                       (int64_t)rows * n->hid);
    gemm<false, true>(n->a2, w + o.W3, w + o.b3, n->y, rows, n->out, n->hid, n->hid, n->hid, n->out, 0);
    LQ_LAUNCH_CHECK();
    return LQ_OK;
}

// backward of the MSE loss on n->y vs n->t; gradients into n->g; loss to n->loss[0]
static int backward(lq_mlp* n, int rows) {
    Off o = offsets(n);
    const float* w = n->w;
    float* g = n->g;
    int64_t ny = (int64_t)rows * n->out;
    unsigned nb = nblk(ny);
    float* partial = n->loss + 1;
    hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, stream(), n->y, n->t, n->d3, ny, 1.0f / (float)ny, partial);
    hipLaunchKernelGGL(k_sum_f, dim3(1), dim3(256), 0, stream(), (const float*)partial, (int)nb, n->loss);
    // layer 3: dW3 = d3^T a2, db3 = colsum(d3), d2 = d3 W3 (then * gelu'(z2))
    gemm<true, false>(n->d3, n->a2, nullptr, g + o.W3, n->out, n->hid, rows, n->out, n->hid, n->hid, 0);
    hipLaunchKernelGGL(k_colsum, dim3(nblk(n->out)), dim3(256), 0, stream(), n->d3, g + o.b3, rows, n->out);
    gemm<false, false>(n->d3, w + o.W3, nullptr, n->d2, rows, n->hid, n->out, n->out, n->hid, n->hid, 0);
    hipLaunchKernelGGL(k_gelu_back, dim3(nblk((int64_t)rows * n->hid)), dim3(256), 0, stream(), n->z2, n->d2,
                       (int64_t)rows * n->hid);
    // layer 2
    gemm<true, false>(n->d2, n->a1, nullptr, g + o.W2, n->hid, n->hid, rows, n->hid, n->hid, n->hid, 0);
    hipLaunchKernelGGL(k_colsum, dim3(nblk(n->hid)), dim3(256), 0, stream(), n->d2, g + o.b2, rows, n->hid);
    gemm<false, false>(n->d2, w + o.W2, nullptr, n->d1, rows, n->hid, n->hid, n->hid, n->hid, n->hid, 0);
    hipLaunchKernelGGL(k_gelu_back, dim3(nblk((int64_t)rows * n->hid)), dim3(256), 0, stream(), n->z1, n->d1,
                       (int64_t)rows * n->hid);
    // layer 1
    hipLaunchKernelGGL(k_layer1_grad, dim3(nblk((int64_t)n->hid * n->in)), dim3(256), 0, stream(), n->x, n->d1,
                       g + o.W1, rows, n->in, n->hid);
    hipLaunchKernelGGL(k_colsum, dim3(nblk(n->hid)), dim3(256), 0, stream(), n->d1, g + o.b1, rows, n->hid);
    LQ_LAUNCH_CHECK();
    return LQ_OK;
}

static int adam(lq_mlp* n, float lr) {
    n->step++;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float c1 = 1.0f / (1.0f - powf(b1, (float)n->step));
    float c2 = 1.0f / (1.0f - powf(b2, (float)n->step));
    hipLaunchKernelGGL(k_adam, dim3(nblk(n->nparams)), dim3(256), 0, stream(), n->w, n->g, n->m, n->v, n->nparams,
                       lr, b1, b2, eps, c1, c2);
    LQ_LAUNCH_CHECK();
    return LQ_OK;
}

// SplitMix64 for the initialisation (uniform in [-a, a]).
static uint64_t sm64(uint64_t& s) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace lq

using namespace lq;

extern "C" {

int lq_mlp_create(int in_dim, int hidden, int out_dim, uint64_t seed, lq_mlp** out) {
    LQ_API_GUARD;
    clear_error();
    if (!out || in_dim < 1 || in_dim > 8 || hidden < 16 || hidden > 4096 || out_dim < 1 || out_dim > 8192) {
        set_error("mlp sizes out of range (in <= 8, 16 <= hidden <= 4096, out <= 8192)");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    int rc = ensure_init();
    if (rc) return rc;
    lq_mlp* n = new lq_mlp();
    n->in = in_dim;
    n->hid = hidden;
    n->out = out_dim;
    n->nparams = (int64_t)hidden * in_dim + hidden + (int64_t)hidden * hidden + hidden + (int64_t)out_dim * hidden + out_dim;
    size_t b = (size_t)n->nparams * 4;
    if (hipMalloc((void**)&n->w, b) != hipSuccess || hipMalloc((void**)&n->g, b) != hipSuccess ||
        hipMalloc((void**)&n->m, b) != hipSuccess || hipMalloc((void**)&n->v, b) != hipSuccess) {
        set_error("mlp allocation failed");
        delete n;
        return LQ_ERR_OUT_OF_MEMORY;
    }
    // uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) for weights and biases
    std::vector<float> h(n->nparams);
    Off o = offsets(n);
    uint64_t s = seed;
    auto fill = [&](int64_t off, int64_t cnt, int fan_in) {
        double a = 1.0 / std::sqrt((double)fan_in);
        for (int64_t i = 0; i < cnt; ++i) {
            double u = (double)(sm64(s) >> 11) * (1.0 / 9007199254740992.0);
            h[off + i] = (float)((2.0 * u - 1.0) * a);
        }
    };
    fill(o.W1, (int64_t)hidden * in_dim, in_dim);
    fill(o.b1, hidden, in_dim);
    fill(o.W2, (int64_t)hidden * hidden, hidden);
    fill(o.b2, hidden, hidden);
    fill(o.W3, (int64_t)out_dim * hidden, hidden);
    fill(o.b3, out_dim, hidden);
    std::vector<float> zero(n->nparams, 0.f);
    if (hipMemcpy(n->w, h.data(), b, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(n->m, zero.data(), b, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(n->v, zero.data(), b, hipMemcpyHostToDevice) != hipSuccess) {
        set_error("mlp upload failed");
        return LQ_ERR_DEVICE;
    }
    *out = n;
    return LQ_OK;
}

int lq_mlp_destroy(lq_mlp* n) {
    LQ_API_GUARD;
    if (!n) return LQ_OK;
    (void)hipStreamSynchronize(stream());
    float* bufs[] = {n->w, n->g, n->m, n->v, n->x, n->t, n->z1, n->a1, n->z2, n->a2, n->y, n->d3, n->d2, n->d1, n->loss};
    for (float* b : bufs)
        if (b) (void)hipFree(b);
    delete n;
    return LQ_OK;
}

int lq_mlp_num_params(const lq_mlp* n, int64_t* count) {
    LQ_API_GUARD;
    if (!n || !count) return LQ_ERR_INVALID_ARGUMENT;
    *count = n->nparams;
    return LQ_OK;
}

int lq_mlp_get_params(const lq_mlp* n, float* out, int64_t count) {
    LQ_API_GUARD;
    clear_error();
    if (!n || !out || count != n->nparams) {
        set_error("parameter buffer must hold exactly %lld floats", (long long)(n ? n->nparams : 0));
        return LQ_ERR_DIMENSION_MISMATCH;
    }
    LQ_HIP_CHECK(hipStreamSynchronize(stream()));
    LQ_HIP_CHECK(hipMemcpy(out, n->w, (size_t)count * 4, hipMemcpyDeviceToHost));
    return LQ_OK;
}

int lq_mlp_set_params(lq_mlp* n, const float* in, int64_t count) {
    LQ_API_GUARD;
    clear_error();
    if (!n || !in || count != n->nparams) {
        set_error("parameter buffer must hold exactly %lld floats", (long long)(n ? n->nparams : 0));
        return LQ_ERR_DIMENSION_MISMATCH;
    }
    LQ_HIP_CHECK(hipStreamSynchronize(stream()));
    LQ_HIP_CHECK(hipMemcpy(n->w, in, (size_t)count * 4, hipMemcpyHostToDevice));
    n->step = 0;
    std::vector<float> zero(count, 0.f);
    LQ_HIP_CHECK(hipMemcpy(n->m, zero.data(), (size_t)count * 4, hipMemcpyHostToDevice));
    LQ_HIP_CHECK(hipMemcpy(n->v, zero.data(), (size_t)count * 4, hipMemcpyHostToDevice));
    return LQ_OK;
}

int lq_mlp_forward(lq_mlp* n, const float* x, int rows, float* y) {
    LQ_API_GUARD;
    clear_error();
    if (!n || !x || !y || rows < 1) {
        set_error("invalid arguments");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    int rc = ensure_cap(n, rows);
    if (rc) return rc;
    LQ_HIP_CHECK(hipMemcpyAsync(n->x, x, (size_t)rows * n->in * 4, hipMemcpyHostToDevice, stream()));
    rc = forward(n, rows);
    if (rc) return rc;
    LQ_HIP_CHECK(hipMemcpyAsync(y, n->y, (size_t)rows * n->out * 4, hipMemcpyDeviceToHost, stream()));
    LQ_HIP_CHECK(hipStreamSynchronize(stream()));
    return LQ_OK;
}

// Loss and parameter gradients of the mean squared error (no update).
int lq_mlp_gradients(lq_mlp* n, const float* x, const float* target, int rows, float* loss, float* grad) {
    LQ_API_GUARD;
    clear_error();
    if (!n || !x || !target || rows < 1) {
        set_error("invalid arguments");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    int rc = ensure_cap(n, rows);
    if (rc) return rc;
    LQ_HIP_CHECK(hipMemcpyAsync(n->x, x, (size_t)rows * n->in * 4, hipMemcpyHostToDevice, stream()));
    LQ_HIP_CHECK(hipMemcpyAsync(n->t, target, (size_t)rows * n->out * 4, hipMemcpyHostToDevice, stream()));
    rc = forward(n, rows);
    if (!rc) rc = backward(n, rows);
    if (rc) return rc;
    if (loss) LQ_HIP_CHECK(hipMemcpyAsync(loss, n->loss, 4, hipMemcpyDeviceToHost, stream()));
    if (grad) LQ_HIP_CHECK(hipMemcpyAsync(grad, n->g, (size_t)n->nparams * 4, hipMemcpyDeviceToHost, stream()));
    LQ_HIP_CHECK(hipStreamSynchronize(stream()));
    return LQ_OK;
}

// Train with Adam on a device-resident data set: `epochs` passes over
// `count` samples in minibatches of `batch` (order reshuffled each epoch with
// SplitMix64(seed)); loss_history (nullable) gets the mean loss per epoch.
int lq_mlp_train(lq_mlp* n, const float* x, const float* target, int count, int batch, int epochs, float lr,
                 uint64_t seed, float* loss_history) {
    LQ_API_GUARD;
    clear_error();
    if (!n || !x || !target || count < 1 || batch < 1 || epochs < 0) {
        set_error("invalid arguments");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    int rc = ensure_cap(n, batch);
    if (rc) return rc;
    std::vector<int> idx(count);
    for (int i = 0; i < count; ++i) idx[i] = i;
    std::vector<float> bx((size_t)batch * n->in), bt((size_t)batch * n->out);
    uint64_t s = seed;
    for (int ep = 0; ep < epochs; ++ep) {
        for (int i = count - 1; i > 0; --i) {
            int j = (int)(sm64(s) % (uint64_t)(i + 1));
            std::swap(idx[i], idx[j]);
        }
        double ep_loss = 0.0;
        int nb = 0;
        for (int b0 = 0; b0 + batch <= count || (b0 == 0 && count < batch); b0 += batch) {
            int rows = std::min(batch, count - b0);
            for (int r = 0; r < rows; ++r) {
                std::memcpy(&bx[(size_t)r * n->in], x + (size_t)idx[b0 + r] * n->in, n->in * 4);
                std::memcpy(&bt[(size_t)r * n->out], target + (size_t)idx[b0 + r] * n->out, n->out * 4);
            }
            LQ_HIP_CHECK(hipMemcpyAsync(n->x, bx.data(), (size_t)rows * n->in * 4, hipMemcpyHostToDevice, stream()));
            LQ_HIP_CHECK(hipMemcpyAsync(n->t, bt.data(), (size_t)rows * n->out * 4, hipMemcpyHostToDevice, stream()));
            rc = forward(n, rows);
            if (!rc) rc = backward(n, rows);
            if (!rc) rc = adam(n, lr);
            if (rc) return rc;
            float l = 0.f;
            LQ_HIP_CHECK(hipMemcpyAsync(&l, n->loss, 4, hipMemcpyDeviceToHost, stream()));
            LQ_HIP_CHECK(hipStreamSynchronize(stream()));
            ep_loss += l;
            ++nb;
            if (count < batch) break;
        }
        if (loss_history) loss_history[ep] = (float)(ep_loss / std::max(nb, 1));
    }
    return LQ_OK;
}

}  // extern "C"
