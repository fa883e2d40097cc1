// Adjoint-method gradient of E(theta) = <psi|H|psi>, psi = U_G ... U_1 |0...0>.
// No counterpart in the reference (it offers ParameterShift and
// FiniteDifference, src/optimization/gradient.rs): one forward sweep, then
// lambda = H psi and a backward sweep.  For a rotation U_k = exp(-i theta/2 P)
// (RX, RY, RZ) dE/dtheta_k = Im <lambda_k| P |psi_k> with both states taken
// right after gate k; then U_k^dagger is applied to psi and lambda.  Equal to
// ParameterShift with shift pi/2 up to rounding.  H uses the same Y-phase
// convention as the expectation (reference or physical).
//
// The backward sweep is fused like the forward one: the inverse circuit
// U_1^dagger ... U_G^dagger is planned into tile passes, and one pass loads
// the psi and lambda tiles together into LDS, applies the pass's ops to both
// and, right before the inverse of a parameterised rotation, adds the tile's
// share of Im <lambda|P|psi> (f64) into a per-(gradient, tile, wavefront)
// partial; a last kernel sums the partials in a fixed order.  The planner's
// reordering of commuting gates is harmless: any order of the factors that is
// the same operator for every angle gives the same derivative.  HBM traffic
// per pass: read and write both states once.
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "lq_fused.h"

namespace lq {

void bind_template(const lq_gate* tmpl, int num_gates, const double* params, int num_params,
                   std::vector<lq_gate>& out);
int check_template(int n, const lq_gate* tmpl, int num_gates, int num_params);
int execute_fused(lq_state* st, const lq_gate* gates, int count);

constexpr int kAThreads = 256;
constexpr int kAWaves = kAThreads / 64;

template <typename V>
__global__ void k_zero(V* __restrict__ a, uint64_t dim) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
        V z;
        z.x = 0;
        z.y = 0;
        a[i] = z;
    }
}

// lambda_i += w_x(i) psi_{i ^ x}, w_x(i) = sum_t (cr_t, ci_t) (-1)^{popc(i & zy_t)}
// (the matrix element <i|H|i^x> of the group's terms).
template <typename V>
__global__ void __launch_bounds__(kAThreads) k_pauli_apply_add(const V* __restrict__ psi, V* __restrict__ lam,
                                                               uint64_t dim, uint64_t x,
                                                               const PauliTermDev* __restrict__ terms, int nt) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < dim; i += stride) {
        double wr = 0.0, wi = 0.0;
        for (int t = 0; t < nt; ++t) {
            bool s = __popcll(i & terms[t].zy) & 1;
            wr += s ? -terms[t].cr : terms[t].cr;
            wi += s ? -terms[t].ci : terms[t].ci;
        }
        V b = psi[i ^ x];
        V l = lam[i];
        double br = (double)b.x, bi = (double)b.y;
        l.x = (decltype(l.x))((double)l.x + (wr * br - wi * bi));
        l.y = (decltype(l.y))((double)l.y + (wr * bi + wi * br));
        lam[i] = l;
    }
}

// Per-op gradient bookkeeping of the backward plan.
struct AdjMeta {
    int32_t slot;   // gradient slot (-1: the op carries no parameter)
    int32_t pauli;  // generator: LQ_PAULI_X / _Y / _Z
    int32_t lb;     // X / Y: local tile bit of the generator
    int32_t gbit;   // Z: global bit of the generator
};

__device__ __forceinline__ uint32_t ins0a(uint32_t k, int p) {
    return ((k >> p) << (p + 1)) | (k & ((1u << p) - 1u));
}

// One op of a tile pass on a tile in LDS (the op-by-op form of the forward
// tile kernel).  Elements: thread-strided, `gidx` maps an element to its
// global amplitude index.
template <int K, typename V, typename T, typename G>
__device__ __forceinline__ void adj_apply(V* tile, const TileOp& op, const G& gidx) {
    constexpr uint32_t TS = 1u << K;
    constexpr int NT = kAThreads;
    constexpr bool FULL = TS >= (uint32_t)NT;
    constexpr int EPT = FULL ? (int)(TS / NT) : 1;
    constexpr bool PFULL = TS / 2 >= (uint32_t)NT;
    constexpr int PPT = PFULL ? (int)(TS / 2 / NT) : 1;
    const uint32_t tid = threadIdx.x;
    const int kind = op.kind;
    if (kind == kTDiag) {
        const uint64_t ctrl = op.ctrl;
        const int s0 = op.sel0, s1 = op.sel1;
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
            const uint32_t e = tid + (uint32_t)j * NT;
            if (!FULL && e >= TS) continue;
            const uint64_t g = gidx(e);
            if ((g & ctrl) != ctrl) continue;
            const int b0 = s0 >= 0 ? (int)((g >> s0) & 1ull) : 0;
            const int b1 = s1 >= 0 ? (int)((g >> s1) & 1ull) : 0;
            const int k = b0 | (b1 << 1);
            tile[e] = cmul((T)op.m[2 * k], (T)op.m[2 * k + 1], tile[e]);
        }
    } else if (kind <= kTPairH) {
        const int lb = op.lb0;
        const uint64_t ctrl = op.ctrl;
        const T m0 = (T)op.m[0], m1 = (T)op.m[1], m2 = (T)op.m[2], m3 = (T)op.m[3];
        const T m4 = (T)op.m[4], m5 = (T)op.m[5], m6 = (T)op.m[6], m7 = (T)op.m[7];
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            const uint32_t q = tid + (uint32_t)j * NT;
            if (!PFULL && q >= TS / 2) continue;
            const uint32_t e0 = ins0a(q, lb);
            const uint32_t e1 = e0 | (1u << lb);
            if (ctrl && (gidx(e0) & ctrl) != ctrl) continue;
            V a0 = tile[e0], a1 = tile[e1];
            V o0, o1;
            if (kind == kTPairMat) {
                o0 = cadd(cmul(m0, m1, a0), cmul(m2, m3, a1));
                o1 = cadd(cmul(m4, m5, a0), cmul(m6, m7, a1));
            } else if (kind == kTPairSwap) {
                o0 = a1;
                o1 = a0;
            } else if (kind == kTPairNegY) {
                o0 = cmul((T)0, (T)1, a1);
                o1 = cmul((T)0, (T)-1, a0);
            } else {
                V sm, df;
                sm.x = a0.x + a1.x;
                sm.y = a0.y + a1.y;
                df.x = a0.x - a1.x;
                df.y = a0.y - a1.y;
                o0 = cmul(m0, (T)0, sm);
                o1 = cmul(m0, (T)0, df);
            }
            tile[e0] = o0;
            tile[e1] = o1;
        }
    } else if (kind == kTSwap2) {
        const int lb0 = op.lb0, lb1 = op.lb1;
        const int lo = min(lb0, lb1), hi = max(lb0, lb1);
        for (uint32_t q = tid; q < TS / 4; q += NT) {
            const uint32_t e = ins0a(ins0a(q, lo), hi) | (1u << lb0);
            const uint32_t f = e ^ (1u << lb0) ^ (1u << lb1);
            V x = tile[e];
            tile[e] = tile[f];
            tile[f] = x;
        }
    } else {  // kTQuad
        const int lo = min(op.lb0, op.lb1), hi = max(op.lb0, op.lb1);
        const uint32_t ma = 1u << op.lb0, mb = 1u << op.lb1;
        for (uint32_t q = tid; q < TS / 4; q += NT) {
            const uint32_t e = ins0a(ins0a(q, lo), hi);
            V a[4] = {tile[e], tile[e | mb], tile[e | ma], tile[e | ma | mb]};
            V y[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                V acc = cmul((T)op.m[8 * r], (T)op.m[8 * r + 1], a[0]);
#pragma unroll
                for (int c = 1; c < 4; ++c)
                    acc = cadd(acc, cmul((T)op.m[8 * r + 2 * c], (T)op.m[8 * r + 2 * c + 1], a[c]));
                y[r] = acc;
            }
            tile[e] = y[0];
            tile[e | mb] = y[1];
            tile[e | ma] = y[2];
            tile[e | ma | mb] = y[3];
        }
    }
}

// Backward tile pass over `ninst` (psi, lambda) pairs.  partial layout:
// [inst][slot][tile][wave] (nwp = tiles * waves per slot).
template <int K, typename V, typename T>
__global__ void __launch_bounds__(kAThreads) k_tile_adj(V* __restrict__ psi, V* __restrict__ lam, TilePass p,
                                                        const TileOp* __restrict__ ops,
                                                        const AdjMeta* __restrict__ meta, uint64_t ninst,
                                                        uint64_t inst_stride, uint64_t ops_stride,
                                                        double* __restrict__ partial, int nslots, uint64_t nwp) {
    constexpr uint32_t TS = 1u << K;
    constexpr int NT = kAThreads;
    constexpr bool FULL = TS >= (uint32_t)NT;
    constexpr int EPT = FULL ? (int)(TS / NT) : 1;
    constexpr bool PFULL = TS / 2 >= (uint32_t)NT;
    constexpr int PPT = PFULL ? (int)(TS / 2 / NT) : 1;
    extern __shared__ __align__(16) unsigned char smem[];
    V* ta = reinterpret_cast<V*>(smem);
    V* tb = ta + TS;
    uint64_t* hoff = reinterpret_cast<uint64_t*>(smem + 2 * TS * sizeof(V));
    const int L = p.L;
    const uint32_t lowmask = (1u << L) - 1u;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    const bool own = FULL || tid < TS;
    for (uint32_t h = tid; h < (1u << p.hb); h += NT) {
        uint64_t o = 0;
        for (int j = 0; j < p.hb; ++j)
            if ((h >> j) & 1u) o |= 1ull << p.hpos[j];
        hoff[h] = o;
    }
    __syncthreads();
    const AdjMeta* md = meta + p.op_first;
    const uint64_t total = ninst * p.ntiles;
    for (uint64_t w = blockIdx.x; w < total; w += gridDim.x) {
        const uint32_t inst = __builtin_amdgcn_readfirstlane((uint32_t)(w / p.ntiles));
        const uint64_t t = w - (uint64_t)inst * p.ntiles;
        V* sa = psi + (uint64_t)inst * inst_stride;
        V* sb = lam + (uint64_t)inst * inst_stride;
        const TileOp* op = ops + (uint64_t)inst * ops_stride + p.op_first;
        uint64_t base = t << L;
        for (int j = 0; j < p.hb; ++j) base = insert_zero(base, p.hpos[j]);
        auto gidx = [&](uint32_t e) -> uint64_t { return base | (e & lowmask) | hoff[e >> L]; };
        if (own) {
            V ra[EPT], rb[EPT];
#pragma unroll
            for (int j = 0; j < EPT; ++j) {
                const uint64_t g = gidx(tid + (uint32_t)j * NT);
                ra[j] = sa[g];
                rb[j] = sb[g];
            }
#pragma unroll
            for (int j = 0; j < EPT; ++j) {
                ta[tid + (uint32_t)j * NT] = ra[j];
                tb[tid + (uint32_t)j * NT] = rb[j];
            }
        }
        __syncthreads();
        for (int o = 0; o < p.op_count; ++o) {
            const int slot = md[o].slot;
            if (slot >= 0) {
                // this tile's share of Im <lambda| P |psi>, before the inverse
                const int pauli = md[o].pauli;
                double acc = 0.0;
                if (pauli == LQ_PAULI_Z) {
                    const int gb = md[o].gbit;
#pragma unroll
                    for (int j = 0; j < EPT; ++j) {
                        const uint32_t e = tid + (uint32_t)j * NT;
                        if (!own) continue;
                        const V l = tb[e], s = ta[e];
                        const double v = (double)l.x * (double)s.y - (double)l.y * (double)s.x;
                        acc += ((gidx(e) >> gb) & 1ull) ? -v : v;
                    }
                } else {
                    const int lb = md[o].lb;
#pragma unroll
                    for (int j = 0; j < PPT; ++j) {
                        const uint32_t q = tid + (uint32_t)j * NT;
                        if (!PFULL && q >= TS / 2) continue;
                        const uint32_t e0 = ins0a(q, lb);
                        const uint32_t e1 = e0 | (1u << lb);
                        const V l0 = tb[e0], l1 = tb[e1], s0 = ta[e0], s1 = ta[e1];
                        if (pauli == LQ_PAULI_X) {
                            // (X psi)_e0 = psi_e1, (X psi)_e1 = psi_e0
                            acc += ((double)l0.x * s1.y - (double)l0.y * s1.x) +
                                   ((double)l1.x * s0.y - (double)l1.y * s0.x);
                        } else {
                            // (Y psi)_e0 = -i psi_e1, (Y psi)_e1 = i psi_e0
                            acc += -((double)l0.x * s1.x + (double)l0.y * s1.y) +
                                   ((double)l1.x * s0.x + (double)l1.y * s0.y);
                        }
                    }
                }
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
                if (lane == 0) partial[((uint64_t)inst * nslots + slot) * nwp + t * kAWaves + wave] = acc;
                __syncthreads();
            }
            adj_apply<K, V, T>(ta, op[o], gidx);
            adj_apply<K, V, T>(tb, op[o], gidx);
            __syncthreads();
        }
        if (own) {
            V ra[EPT], rb[EPT];
#pragma unroll
            for (int j = 0; j < EPT; ++j) {
                ra[j] = ta[tid + (uint32_t)j * NT];
                rb[j] = tb[tid + (uint32_t)j * NT];
            }
#pragma unroll
            for (int j = 0; j < EPT; ++j) {
                const uint64_t g = gidx(tid + (uint32_t)j * NT);
                sa[g] = ra[j];
                sb[g] = rb[j];
            }
        }
        __syncthreads();
    }
}

// out[r] = sum_i partial[r * nwp + i], fixed order (one block per row).
__global__ void __launch_bounds__(kAThreads) k_rows_sum(const double* __restrict__ partial, uint64_t nwp,
                                                        double* __restrict__ out) {
    __shared__ double sh[kAThreads];
    const double* row = partial + (uint64_t)blockIdx.x * nwp;
    double s = 0.0;
    for (uint64_t i = threadIdx.x; i < nwp; i += kAThreads) s += row[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int k = kAThreads / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

template <int K, typename V, typename T>
static void launch_adj(void* psi, void* lam, const TilePass& p, const TileOp* ops, const AdjMeta* meta,
                       uint64_t ninst, uint64_t stride, uint64_t ostride, double* partial, int nslots, uint64_t nwp,
                       unsigned grid, size_t lds) {
    static std::once_flag once;
    std::call_once(once, [&] {
        (void)hipFuncSetAttribute((const void*)k_tile_adj<K, V, T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds + 4096);
    });
    hipLaunchKernelGGL((k_tile_adj<K, V, T>), dim3(grid), dim3(kAThreads), lds, stream(), (V*)psi, (V*)lam, p, ops,
                       meta, ninst, stride, ostride, partial, nslots, nwp);
}

template <typename V, typename T>
static int dispatch_adj(int K, void* psi, void* lam, const TilePass& p, const TileOp* ops, const AdjMeta* meta,
                        uint64_t ninst, uint64_t stride, uint64_t ostride, double* partial, int nslots, uint64_t nwp) {
    const size_t lds = 2 * (sizeof(V) << K) + (sizeof(uint64_t) << p.hb);
    const unsigned grid = (unsigned)std::min<uint64_t>(ninst * p.ntiles, 8192);
    switch (K) {
#define LQ_KA(k)                                                                                             \
    case k:                                                                                                  \
        launch_adj<k, V, T>(psi, lam, p, ops, meta, ninst, stride, ostride, partial, nslots, nwp, grid, lds); \
        break;
        LQ_KA(1) LQ_KA(2) LQ_KA(3) LQ_KA(4) LQ_KA(5) LQ_KA(6) LQ_KA(7) LQ_KA(8) LQ_KA(9) LQ_KA(10) LQ_KA(11)
        LQ_KA(12)
#undef LQ_KA
        default:
            set_error("adjoint tile of 2^%d amplitudes is not supported", K);
            return LQ_ERR_UNSUPPORTED;
    }
    LQ_LAUNCH_CHECK();
    return LQ_OK;
}

// Backward-sweep tiles: two tiles (psi, lambda) of 32 KiB each in LDS.
static int adj_tile_bits(int n, int dtype) { return std::min(n, dtype == LQ_C128 ? 11 : 12); }

// Device buffers of the gradient (grown on demand, reused): the backward
// plan's ops and metadata, Pauli terms, partials and results.  Dedicated
// because the expectation and normalisation helpers use the shared scratch.
static void* g_adj = nullptr;
static size_t g_adj_cap = 0;
static unsigned char* g_adj_host = nullptr;
static size_t g_adj_host_cap = 0;

static char* adj_buffer(size_t bytes) {
    if (bytes <= g_adj_cap) return (char*)g_adj;
    if (hipStreamSynchronize(stream()) != hipSuccess) return nullptr;
    if (g_adj) (void)hipFree(g_adj);
    g_adj = nullptr;
    g_adj_cap = 0;
    if (hipMalloc(&g_adj, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    g_adj_cap = bytes;
    return (char*)g_adj;
}

static unsigned char* adj_host(size_t bytes) {
    if (bytes <= g_adj_host_cap) return g_adj_host;
    if (hipStreamSynchronize(stream()) != hipSuccess) return nullptr;
    if (g_adj_host) (void)hipHostFree(g_adj_host);
    g_adj_host = nullptr;
    g_adj_host_cap = 0;
    if (hipHostMalloc((void**)&g_adj_host, bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    g_adj_host_cap = bytes;
    return g_adj_host;
}

// Cached psi / lambda states (same n and dtype across VQE iterations).
static lq_state* g_adj_psi = nullptr;
static lq_state* g_adj_lam = nullptr;

static int adj_states(int n, int dtype, lq_state** psi, lq_state** lam) {
    if (!g_adj_psi || g_adj_psi->n != n || g_adj_psi->dtype != dtype) {
        if (g_adj_psi) lq_state_destroy(g_adj_psi);
        if (g_adj_lam) lq_state_destroy(g_adj_lam);
        g_adj_psi = g_adj_lam = nullptr;
        int rc = lq_state_create(n, dtype, &g_adj_psi);
        if (rc) return rc;
        rc = lq_state_create(n, dtype, &g_adj_lam);
        if (rc) {
            lq_state_destroy(g_adj_psi);
            g_adj_psi = nullptr;
            return rc;
        }
    }
    *psi = g_adj_psi;
    *lam = g_adj_lam;
    return LQ_OK;
}

static unsigned grid_of(uint64_t work) {
    uint64_t g = (work + kAThreads - 1) / kAThreads;
    if (g > 8192) g = 8192;
    if (g < 1) g = 1;
    return (unsigned)g;
}

// Inverse of a gate as a gate descriptor; false if not supported.
static bool inverse_gate(const lq_gate& g, lq_gate& inv) {
    inv = g;
    switch (g.kind) {
        case LQ_GATE_X: case LQ_GATE_Y: case LQ_GATE_Z: case LQ_GATE_H: case LQ_GATE_CNOT: case LQ_GATE_SWAP:
        case LQ_GATE_CZ: case LQ_GATE_TOFFOLI:
            return true;  // involutions (the reference's -Y squares to I too)
        case LQ_GATE_S:
            inv.kind = LQ_GATE_PHASE;
            inv.theta[0] = -M_PI / 2.0;
            return true;
        case LQ_GATE_T:
            inv.kind = LQ_GATE_PHASE;
            inv.theta[0] = -M_PI / 4.0;
            return true;
        case LQ_GATE_RX: case LQ_GATE_RY: case LQ_GATE_RZ: case LQ_GATE_PHASE: case LQ_GATE_CPHASE:
        case LQ_GATE_CRX: case LQ_GATE_CRY: case LQ_GATE_CRZ: case LQ_GATE_CP: case LQ_GATE_RXX: case LQ_GATE_RYY:
        case LQ_GATE_RZZ:
            inv.theta[0] = -g.theta[0];
            return true;
        case LQ_GATE_U3:
            // U3(t, p, l)^-1 = U3(-t, -l, -p)
            inv.theta[0] = -g.theta[0];
            inv.theta[1] = -g.theta[2];
            inv.theta[2] = -g.theta[1];
            return true;
        default:
            return false;
    }
}

template <typename V, typename T>
static int adjoint_t(int n, int dtype, const lq_gate* tmpl, int num_gates, const double* params, int num_params,
                     const PauliSum& ps, double* energy_out, double* grad_out) {
    std::vector<lq_gate> g;
    bind_template(tmpl, num_gates, params, num_params, g);
    // the backward circuit: U_G^dagger first
    std::vector<lq_gate> rev(num_gates);
    std::vector<int> slot_of(num_gates, -1);
    int nslots = 0;
    for (int k = 0; k < num_gates; ++k) {
        if ((g[k].flags & LQ_GATE_FLAG_RENORMALIZE) || !inverse_gate(g[k], rev[num_gates - 1 - k])) {
            set_error("adjoint gradient: gate %d (kind %d) has no unitary inverse here", k, g[k].kind);
            return LQ_ERR_UNSUPPORTED;
        }
        for (int a = 0; a < 3; ++a)
            if (tmpl[k].param[a] >= 0 && !(a == 0 && (g[k].kind == LQ_GATE_RX || g[k].kind == LQ_GATE_RY ||
                                                      g[k].kind == LQ_GATE_RZ))) {
                set_error("adjoint gradient supports parameters on RX / RY / RZ only (gate %d)", k);
                return LQ_ERR_UNSUPPORTED;
            }
        if (tmpl[k].param[0] >= 0) slot_of[k] = nslots++;
    }
    std::vector<Prim> prims;
    int rc = gates_to_prims(n, rev.data(), num_gates, prims);
    if (rc) return rc;
    const int K = adj_tile_bits(n, dtype);
    const int L = std::min(default_low_bits(n, dtype), K);
    Plan plan;
    plan_passes(n, prims, K, L, plan);
    std::vector<AdjMeta> meta(plan.ops.size());
    for (size_t o = 0; o < plan.ops.size(); ++o) {
        // the planner's grouping step relabels 2x2 ops by matrix class for the
        // grouped forward kernel; the backward kernel applies every 2x2 op in
        // the general form (the matrix is complete in m[0..7])
        TileOp& op = plan.ops[o];
        if (op.kind == kTPairReal || op.kind == kTPairRX) op.kind = kTPairMat;
        if (op.kind < kTPairMat || op.kind > kTQuad) {
            set_error("adjoint plan: unexpected op kind %d", op.kind);
            return LQ_ERR_DEVICE;
        }
        const int k = num_gates - 1 - plan.op_src[o];
        AdjMeta m{-1, 0, 0, 0};
        if (k >= 0 && k < num_gates && slot_of[k] >= 0) {
            m.slot = slot_of[k];
            m.pauli = g[k].kind == LQ_GATE_RX ? LQ_PAULI_X : g[k].kind == LQ_GATE_RY ? LQ_PAULI_Y : LQ_PAULI_Z;
            m.gbit = n - 1 - g[k].qubits[0];
            m.lb = plan.ops[o].lb0;
            if (m.pauli != LQ_PAULI_Z && m.lb < 0) {
                set_error("adjoint plan: rotation %d is not on a tile bit", k);
                return LQ_ERR_DEVICE;
            }
        }
        meta[o] = m;
    }
    lq_state *psi = nullptr, *lam = nullptr;
    rc = adj_states(n, dtype, &psi, &lam);
    if (rc) return rc;
    const uint64_t dim = 1ull << n;
    const uint64_t ntiles = dim >> K;
    const uint64_t nwp = ntiles * kAWaves;
    // device layout: [ops][meta][terms][partials][results]
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t ob = al(plan.ops.size() * sizeof(TileOp) + 1);
    const size_t mb = al(meta.size() * sizeof(AdjMeta) + 1);
    const size_t tb = al(std::max<size_t>(ps.terms.size(), 1) * sizeof(PauliTermDev));
    const size_t pb = al((size_t)std::max(nslots, 1) * nwp * sizeof(double));
    const size_t rb = al((size_t)std::max(nslots, 1) * sizeof(double));
    char* base = adj_buffer(ob + mb + tb + pb + rb);
    unsigned char* hbase = adj_host(ob + mb + tb + rb);
    if (!base || !hbase) {
        set_error("adjoint buffer allocation failed");
        return LQ_ERR_OUT_OF_MEMORY;
    }
    TileOp* d_ops = (TileOp*)base;
    AdjMeta* d_meta = (AdjMeta*)(base + ob);
    PauliTermDev* d_terms = (PauliTermDev*)(base + ob + mb);
    double* d_part = (double*)(base + ob + mb + tb);
    double* d_res = (double*)(base + ob + mb + tb + pb);
    // staging: the host buffer may still feed the previous call's copy
    LQ_HIP_CHECK(hipStreamSynchronize(stream()));
    if (!plan.ops.empty()) std::memcpy(hbase, plan.ops.data(), plan.ops.size() * sizeof(TileOp));
    if (!meta.empty()) std::memcpy(hbase + ob, meta.data(), meta.size() * sizeof(AdjMeta));
    if (!ps.terms.empty()) std::memcpy(hbase + ob + mb, ps.terms.data(), ps.terms.size() * sizeof(PauliTermDev));
    LQ_HIP_CHECK(hipMemcpyAsync(base, hbase, ob + mb + tb, hipMemcpyHostToDevice, stream()));
    // forward sweep (fused tile passes), then lambda = H psi
    rc = execute_fused(psi, g.data(), num_gates);
    if (rc) return rc;
    if (energy_out) {
        rc = expect_pauli_sum(psi, ps, energy_out);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_zero<V>, dim3(grid_of(dim)), dim3(kAThreads), 0, stream(), (V*)lam->d_amps, dim);
    for (const PauliGroup& grp : ps.groups)
        hipLaunchKernelGGL(k_pauli_apply_add<V>, dim3(grid_of(dim)), dim3(kAThreads), 0, stream(),
                           (const V*)psi->d_amps, (V*)lam->d_amps, dim, grp.x,
                           (const PauliTermDev*)(d_terms + grp.first), grp.count);
    LQ_LAUNCH_CHECK();
    // backward sweep
    for (const Step& s : plan.steps) {
        if (s.normalize) continue;  // no renormalising gates reach here
        rc = dispatch_adj<V, T>(K, psi->d_amps, lam->d_amps, s.pass, d_ops, d_meta, 1, 0, 0, d_part, nslots, nwp);
        if (rc) return rc;
    }
    if (nslots > 0) {
        hipLaunchKernelGGL(k_rows_sum, dim3(nslots), dim3(kAThreads), 0, stream(), (const double*)d_part, nwp,
                           d_res);
        LQ_LAUNCH_CHECK();
    }
    double* hres = (double*)(hbase + ob + mb + tb);
    if (nslots > 0)
        LQ_HIP_CHECK(hipMemcpyAsync(hres, d_res, nslots * sizeof(double), hipMemcpyDeviceToHost, stream()));
    LQ_HIP_CHECK(hipStreamSynchronize(stream()));
    for (int i = 0; i < num_params; ++i) grad_out[i] = 0.0;
    for (int k = 0; k < num_gates; ++k)
        if (slot_of[k] >= 0) grad_out[tmpl[k].param[0]] += hres[slot_of[k]];
    return LQ_OK;
}

}  // namespace lq

using namespace lq;

extern "C" int lq_adjoint_gradient(int num_qubits, int dtype, const lq_gate* tmpl, int num_gates,
                                   const double* params, int num_params, const uint8_t* paulis, const double* coeffs,
                                   int num_terms, int convention, double* energy_out, double* grad_out) {
    clear_error();
    if (num_qubits < 1 || num_qubits > kMaxQubits || num_params < 0 || (num_params > 0 && (!params || !grad_out)) ||
        (dtype != LQ_C128 && dtype != LQ_C64)) {
        set_error("invalid arguments");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    int rc = check_template(num_qubits, tmpl, num_gates, num_params);
    if (rc) return rc;
    PauliSum ps;
    rc = compile_pauli_sum(num_qubits, paulis, coeffs, num_terms, convention, &ps);
    if (rc) return rc;
    if (dtype == LQ_C128)
        return adjoint_t<double2, double>(num_qubits, dtype, tmpl, num_gates, params, num_params, ps, energy_out,
                                          grad_out);
    return adjoint_t<float2, float>(num_qubits, dtype, tmpl, num_gates, params, num_params, ps, energy_out,
                                    grad_out);
}
