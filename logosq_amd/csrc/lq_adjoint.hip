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
constexpr int kAThreads = 256;
constexpr int kAWaves = kAThreads / 64;

// States [count][dim]: amplitude 0 of every instance 1 (|0...0>), or all 0.
template <typename V>
__global__ void k_init_states(V* __restrict__ a, uint64_t dim, uint64_t total, int basis0) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        V z;
        z.x = (basis0 && (i & (dim - 1)) == 0) ? 1 : 0;
        z.y = 0;
        a[i] = z;
    }
}

// lambda_i += w_x(i) psi_{i ^ x}, w_x(i) = sum_t (cr_t, ci_t) (-1)^{popc(i & zy_t)}
// (the matrix element <i|H|i^x> of the group's terms), for every instance:
// element e of the batch is amplitude e & (dim - 1) of instance e >> n, whose
// terms start at terms + instance * tstride.
template <typename V>
__global__ void __launch_bounds__(kAThreads) k_pauli_apply_add(const V* __restrict__ psi, V* __restrict__ lam,
                                                               int n, uint64_t total, uint64_t x,
                                                               const PauliTermDev* __restrict__ terms,
                                                               int tstride, int nt) {
    const uint64_t dmask = (1ull << n) - 1ull;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const uint64_t i = e & dmask;
        const PauliTermDev* tt = terms + (e >> n) * (uint64_t)tstride;
        double wr = 0.0, wi = 0.0;
        for (int t = 0; t < nt; ++t) {
            bool s = __popcll(i & tt[t].zy) & 1;
            wr += s ? -tt[t].cr : tt[t].cr;
            wi += s ? -tt[t].ci : tt[t].ci;
        }
        V b = psi[e ^ x];
        V l = lam[e];
        double br = (double)b.x, bi = (double)b.y;
        l.x = (decltype(l.x))((double)l.x + (wr * br - wi * bi));
        l.y = (decltype(l.y))((double)l.y + (wr * bi + wi * br));
        lam[e] = l;
    }
}

// partial[inst * gridDim.x + b] = block b's share of Re <psi_inst | lambda_inst>
// (blockIdx.y = instance); summed by k_rows_sum in a fixed order.
template <typename V>
__global__ void __launch_bounds__(kAThreads) k_re_inner(const V* __restrict__ psi, const V* __restrict__ lam,
                                                        uint64_t dim, double* __restrict__ partial) {
    __shared__ double sh[kAThreads];
    const V* a = psi + (uint64_t)blockIdx.y * dim;
    const V* b = lam + (uint64_t)blockIdx.y * dim;
    double s = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * kAThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kAThreads + threadIdx.x; i < dim; i += stride)
        s += (double)a[i].x * (double)b[i].x + (double)a[i].y * (double)b[i].y;
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int k = kAThreads / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(uint64_t)blockIdx.y * gridDim.x + blockIdx.x] = sh[0];
}

// Per-op gradient bookkeeping of the backward plan.
struct AdjMeta {
    int32_t slot;   // gradient slot (-1: the op carries no parameter)
    int32_t pauli;  // generator: LQ_PAULI_X / _Y / _Z
    int32_t lb;     // X / Y: local tile bit of the generator
    int32_t gbit;   // Z: global bit of the generator
};

// A run of consecutive ops of a backward pass applied in one sweep over the
// tile: every 2x2 op of the run acts on local bit lb (diagonal ops may act
// anywhere), and at most kRunParams of them carry a parameter.  lb = -1: a
// single SWAP / 4x4 op, applied op by op.
constexpr int kRunParams = 4;
struct AdjRun {
    int32_t first;  // first op, relative to the pass
    int32_t count;
    int32_t lb;
    int32_t pad;
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

// on ? x : y, component by component (a select of the whole struct would
// keep the arrays it lands in out of registers)
template <typename V>
__device__ __forceinline__ V vsel(bool on, V x, V y) {
    V r;
    r.x = on ? x.x : y.x;
    r.y = on ? x.y : y.y;
    return r;
}

// The fields of an op (and its gradient bookkeeping) that a run reads.
struct OpLite {
    int kind, sel0, sel1, slot, pauli, gbit;
    uint64_t ctrl;
    double m[8];
};

__device__ __forceinline__ OpLite op_lite(const TileOp& t, const AdjMeta& md) {
    OpLite r;
    r.kind = t.kind;
    r.sel0 = t.sel0;
    r.sel1 = t.sel1;
    r.ctrl = t.ctrl;
#pragma unroll
    for (int i = 0; i < 8; ++i) r.m[i] = t.m[i];
    r.slot = md.slot;
    r.pauli = md.pauli;
    r.gbit = md.gbit;
    return r;
}

// Backward tile pass over `ninst` (psi, lambda) pairs.  partial layout:
// [inst][slot][tile][wave] (nwp = tiles * waves per slot).
template <int K, typename V, typename T>
__global__ void __launch_bounds__(kAThreads) k_tile_adj(V* __restrict__ psi, V* __restrict__ lam, TilePass p,
                                                        const TileOp* __restrict__ ops,
                                                        const AdjMeta* __restrict__ meta,
                                                        const AdjRun* __restrict__ runs, int nruns, uint64_t ninst,
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
        for (int r = 0; r < nruns; ++r) {
            const int rfirst = runs[r].first, rcount = runs[r].count, lb = runs[r].lb;
            if (lb < 0) {
                adj_apply<K, V, T>(ta, op[rfirst], gidx);
                adj_apply<K, V, T>(tb, op[rfirst], gidx);
                __syncthreads();
                continue;
            }
            const uint64_t gmask = 1ull << (lb < L ? lb : p.hpos[lb - L]);
            double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;
            // this thread's pairs (e0, e1) along bit lb, of both tiles, in registers
            V a0[PPT], a1[PPT], b0[PPT], b1[PPT];
            uint64_t g0[PPT];
            bool live[PPT];
#pragma unroll
            for (int j = 0; j < PPT; ++j) {
                const uint32_t q = tid + (uint32_t)j * NT;
                live[j] = PFULL || q < TS / 2;
                const uint32_t e0 = ins0a(live[j] ? q : 0u, lb);
                const uint32_t e1 = e0 | (1u << lb);
                g0[j] = gidx(e0);
                a0[j] = ta[e0];
                a1[j] = ta[e1];
                b0[j] = tb[e0];
                b1[j] = tb[e1];
            }
            int k = 0;
            // the next op's data is fetched while the current one runs
            OpLite nxt = op_lite(op[rfirst], md[rfirst]);
            for (int o = rfirst; o < rfirst + rcount; ++o) {
                const OpLite cur = nxt;
                if (o + 1 < rfirst + rcount) nxt = op_lite(op[o + 1], md[o + 1]);
                if (cur.slot >= 0) {
                    // share of Im <lambda| P |psi> right before the inverse
                    double c = 0.0;
#pragma unroll
                    for (int j = 0; j < PPT; ++j) {
                        double cj;
                        if (cur.pauli == LQ_PAULI_Z) {
                            const double v0 = (double)b0[j].x * a0[j].y - (double)b0[j].y * a0[j].x;
                            const double v1 = (double)b1[j].x * a1[j].y - (double)b1[j].y * a1[j].x;
                            const uint64_t g1 = g0[j] | gmask;
                            cj = (((g0[j] >> cur.gbit) & 1ull) ? -v0 : v0) + (((g1 >> cur.gbit) & 1ull) ? -v1 : v1);
                        } else if (cur.pauli == LQ_PAULI_X) {
                            // (X psi)_e0 = psi_e1, (X psi)_e1 = psi_e0
                            cj = ((double)b0[j].x * a1[j].y - (double)b0[j].y * a1[j].x) +
                                 ((double)b1[j].x * a0[j].y - (double)b1[j].y * a0[j].x);
                        } else {
                            // (Y psi)_e0 = -i psi_e1, (Y psi)_e1 = i psi_e0
                            cj = -((double)b0[j].x * a1[j].x + (double)b0[j].y * a1[j].y) +
                                 ((double)b1[j].x * a0[j].x + (double)b1[j].y * a0[j].y);
                        }
                        c += live[j] ? cj : 0.0;
                    }
                    acc0 += k == 0 ? c : 0.0;
                    acc1 += k == 1 ? c : 0.0;
                    acc2 += k == 2 ? c : 0.0;
                    acc3 += k == 3 ? c : 0.0;
                    ++k;
                }
                const uint64_t ctrl = cur.ctrl;
                const T m0 = (T)cur.m[0], m1 = (T)cur.m[1], m2 = (T)cur.m[2], m3 = (T)cur.m[3];
                const T m4 = (T)cur.m[4], m5 = (T)cur.m[5], m6 = (T)cur.m[6], m7 = (T)cur.m[7];
                if (cur.kind == kTDiag) {
                    const int s0 = cur.sel0, s1 = cur.sel1;
#pragma unroll
                    for (int j = 0; j < PPT; ++j) {
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            const uint64_t g = h ? (g0[j] | gmask) : g0[j];
                            const bool on = (g & ctrl) == ctrl;
                            const bool c0 = s0 >= 0 && ((g >> s0) & 1ull);
                            const bool c1 = s1 >= 0 && ((g >> s1) & 1ull);
                            const T pr = on ? (c1 ? (c0 ? m6 : m4) : (c0 ? m2 : m0)) : (T)1;
                            const T pi = on ? (c1 ? (c0 ? m7 : m5) : (c0 ? m3 : m1)) : (T)0;
                            if (h) {
                                a1[j] = cmul(pr, pi, a1[j]);
                                b1[j] = cmul(pr, pi, b1[j]);
                            } else {
                                a0[j] = cmul(pr, pi, a0[j]);
                                b0[j] = cmul(pr, pi, b0[j]);
                            }
                        }
                    }
                } else if (cur.kind == kTPairMat) {
                    // 2x2 on bit lb (controls never include lb); lanes whose
                    // controls are not all 1 keep their values
#pragma unroll
                    for (int j = 0; j < PPT; ++j) {
                        const bool on = (g0[j] & ctrl) == ctrl;
                        const V x0 = cadd(cmul(m0, m1, a0[j]), cmul(m2, m3, a1[j]));
                        const V x1 = cadd(cmul(m4, m5, a0[j]), cmul(m6, m7, a1[j]));
                        const V y0 = cadd(cmul(m0, m1, b0[j]), cmul(m2, m3, b1[j]));
                        const V y1 = cadd(cmul(m4, m5, b0[j]), cmul(m6, m7, b1[j]));
                        a0[j] = vsel(on, x0, a0[j]);
                        a1[j] = vsel(on, x1, a1[j]);
                        b0[j] = vsel(on, y0, b0[j]);
                        b1[j] = vsel(on, y1, b1[j]);
                    }
                } else if (cur.kind == kTPairSwap) {
#pragma unroll
                    for (int j = 0; j < PPT; ++j) {
                        const bool on = (g0[j] & ctrl) == ctrl;
                        const V x0 = a0[j], y0 = b0[j];
                        a0[j] = vsel(on, a1[j], a0[j]);
                        a1[j] = vsel(on, x0, a1[j]);
                        b0[j] = vsel(on, b1[j], b0[j]);
                        b1[j] = vsel(on, y0, b1[j]);
                    }
                } else if (cur.kind == kTPairNegY) {
#pragma unroll
                    for (int j = 0; j < PPT; ++j) {
                        const bool on = (g0[j] & ctrl) == ctrl;
                        const V x0 = cmul((T)0, (T)1, a1[j]), x1 = cmul((T)0, (T)-1, a0[j]);
                        const V y0 = cmul((T)0, (T)1, b1[j]), y1 = cmul((T)0, (T)-1, b0[j]);
                        a0[j] = vsel(on, x0, a0[j]);
                        a1[j] = vsel(on, x1, a1[j]);
                        b0[j] = vsel(on, y0, b0[j]);
                        b1[j] = vsel(on, y1, b1[j]);
                    }
                } else {  // kTPairH: m0 = 1/sqrt(2)
#pragma unroll
                    for (int j = 0; j < PPT; ++j) {
                        const bool on = (g0[j] & ctrl) == ctrl;
                        V sa, da, sb, db;
                        sa.x = a0[j].x + a1[j].x;
                        sa.y = a0[j].y + a1[j].y;
                        da.x = a0[j].x - a1[j].x;
                        da.y = a0[j].y - a1[j].y;
                        sb.x = b0[j].x + b1[j].x;
                        sb.y = b0[j].y + b1[j].y;
                        db.x = b0[j].x - b1[j].x;
                        db.y = b0[j].y - b1[j].y;
                        a0[j] = vsel(on, cmul(m0, (T)0, sa), a0[j]);
                        a1[j] = vsel(on, cmul(m0, (T)0, da), a1[j]);
                        b0[j] = vsel(on, cmul(m0, (T)0, sb), b0[j]);
                        b1[j] = vsel(on, cmul(m0, (T)0, db), b1[j]);
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < PPT; ++j) {
                if (!live[j]) continue;
                const uint32_t e0 = ins0a(tid + (uint32_t)j * NT, lb);
                const uint32_t e1 = e0 | (1u << lb);
                ta[e0] = a0[j];
                ta[e1] = a1[j];
                tb[e0] = b0[j];
                tb[e1] = b1[j];
            }
            // the run's gradient shares: one partial per wavefront
            k = 0;
            for (int o = rfirst; o < rfirst + rcount; ++o) {
                const int slot = md[o].slot;
                if (slot < 0) continue;
                double acc = k == 0 ? acc0 : k == 1 ? acc1 : k == 2 ? acc2 : acc3;
                ++k;
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
                if (lane == 0) partial[((uint64_t)inst * nslots + slot) * nwp + t * kAWaves + wave] = acc;
            }
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
                       const AdjRun* runs, int nruns, uint64_t ninst, uint64_t stride, uint64_t ostride, double* partial, int nslots, uint64_t nwp,
                       unsigned grid, size_t lds) {
    static std::once_flag once;
    std::call_once(once, [&] {
        (void)hipFuncSetAttribute((const void*)k_tile_adj<K, V, T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds + 4096);
    });
    hipLaunchKernelGGL((k_tile_adj<K, V, T>), dim3(grid), dim3(kAThreads), lds, stream(), (V*)psi, (V*)lam, p, ops,
                       meta, runs, nruns, ninst, stride, ostride, partial, nslots, nwp);
}

template <typename V, typename T>
static int dispatch_adj(int K, void* psi, void* lam, const TilePass& p, const TileOp* ops, const AdjMeta* meta,
                        const AdjRun* runs, int nruns, uint64_t ninst, uint64_t stride, uint64_t ostride, double* partial, int nslots, uint64_t nwp) {
    const size_t lds = 2 * (sizeof(V) << K) + (sizeof(uint64_t) << p.hb);
    const unsigned grid = (unsigned)std::min<uint64_t>(ninst * p.ntiles, 8192);
    switch (K) {
#define LQ_KA(k)                                                                                             \
    case k:                                                                                                  \
        launch_adj<k, V, T>(psi, lam, p, ops, meta, runs, nruns, ninst, stride, ostride, partial, nslots, nwp, \
                            grid, lds);                                                                     \
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

// Backward-sweep tiles: 2^11 amplitudes of psi and of lambda in LDS (64 KiB
// for complex128, 32 KiB for complex64); each thread holds 4 pairs of each
// in registers through a run.
static int adj_tile_bits(int n, int dtype) {
    (void)dtype;
    return std::min(n, 11);
}

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

// Device states of a batch: psi then lambda, [count][dim] each (grown on
// demand, reused across calls).
static void* g_adj_states = nullptr;
static size_t g_adj_states_cap = 0;

static void* adj_states(size_t bytes) {
    if (bytes <= g_adj_states_cap) return g_adj_states;
    if (hipStreamSynchronize(stream()) != hipSuccess) return nullptr;
    if (g_adj_states) (void)hipFree(g_adj_states);
    g_adj_states = nullptr;
    g_adj_states_cap = 0;
    if (hipMalloc(&g_adj_states, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    g_adj_states_cap = bytes;
    return g_adj_states;
}

// First primitive of every gate (each gate here yields exactly one).
static std::vector<int> first_prim_of(const std::vector<Prim>& prims, int num_gates) {
    std::vector<int> f(num_gates, -1);
    for (size_t k = 0; k < prims.size(); ++k)
        if (prims[k].kind >= 0 && f[prims[k].src] < 0) f[prims[k].src] = (int)k;
    return f;
}

// count instances of one template: parameters params[i * num_params ...],
// Hamiltonian coefficients coeffs[i * num_terms ...] over shared Pauli
// strings.  One plan per direction serves every instance (the structure does
// not depend on the angles); each instance gets its own op table.
template <typename V, typename T>
static int adjoint_batch_t(int n, int dtype, const lq_gate* tmpl, int num_gates, int count, const double* params,
                           int num_params, const uint8_t* paulis, const double* coeffs, int num_terms,
                           int convention, double* energies, double* grads) {
    // template-level checks and gradient slots
    std::vector<int> slot_of(num_gates, -1);
    int nslots = 0;
    for (int k = 0; k < num_gates; ++k) {
        lq_gate inv;
        if ((tmpl[k].flags & LQ_GATE_FLAG_RENORMALIZE) || !inverse_gate(tmpl[k], inv)) {
            set_error("adjoint gradient: gate %d (kind %d) has no unitary inverse here", k, tmpl[k].kind);
            return LQ_ERR_UNSUPPORTED;
        }
        for (int a = 0; a < 3; ++a)
            if (tmpl[k].param[a] >= 0 && !(a == 0 && (tmpl[k].kind == LQ_GATE_RX || tmpl[k].kind == LQ_GATE_RY ||
                                                      tmpl[k].kind == LQ_GATE_RZ))) {
                set_error("adjoint gradient supports parameters on RX / RY / RZ only (gate %d)", k);
                return LQ_ERR_UNSUPPORTED;
            }
        if (tmpl[k].param[0] >= 0) slot_of[k] = nslots++;
    }
    // Pauli sums: one structure (the strings), per-instance coefficients
    PauliSum ps0;
    int rc = compile_pauli_sum(n, paulis, coeffs, num_terms, convention, &ps0);
    if (rc) return rc;
    const int nt = (int)ps0.terms.size();
    // plans from instance 0
    std::vector<lq_gate> g, rev(num_gates);
    bind_template(tmpl, num_gates, params, num_params, g);
    for (int k = 0; k < num_gates; ++k) inverse_gate(g[k], rev[num_gates - 1 - k]);
    std::vector<Prim> pf, pb;
    rc = gates_to_prims(n, g.data(), num_gates, pf);
    if (!rc) rc = gates_to_prims(n, rev.data(), num_gates, pb);
    if (rc) return rc;
    const int Kf = default_tile_bits(n, dtype), Lf = default_low_bits(n, dtype);
    const int Kb = adj_tile_bits(n, dtype), Lb = std::min(default_low_bits(n, dtype), Kb);
    Plan plf, plb;
    plan_passes(n, pf, Kf, Lf, plf);
    plan_passes(n, pb, Kb, Lb, plb);
    const std::vector<int> fpf = first_prim_of(pf, num_gates), fpb = first_prim_of(pb, num_gates);
    std::vector<AdjMeta> meta(plb.ops.size());
    for (size_t o = 0; o < plb.ops.size(); ++o) {
        // the planner's grouping step relabels 2x2 ops by matrix class for the
        // grouped forward kernel; the backward kernel applies every 2x2 op in
        // the general form (the matrix is complete in m[0..7])
        TileOp& op = plb.ops[o];
        if (op.kind == kTPairReal || op.kind == kTPairRX) op.kind = kTPairMat;
        if (op.kind < kTPairMat || op.kind > kTQuad) {
            set_error("adjoint plan: unexpected op kind %d", op.kind);
            return LQ_ERR_DEVICE;
        }
        const int k = num_gates - 1 - plb.op_src[o];
        AdjMeta m{-1, 0, 0, 0};
        if (k >= 0 && k < num_gates && slot_of[k] >= 0) {
            m.slot = slot_of[k];
            m.pauli = g[k].kind == LQ_GATE_RX ? LQ_PAULI_X : g[k].kind == LQ_GATE_RY ? LQ_PAULI_Y : LQ_PAULI_Z;
            m.gbit = n - 1 - g[k].qubits[0];
            m.lb = op.lb0;
            if (m.pauli != LQ_PAULI_Z && m.lb < 0) {
                set_error("adjoint plan: rotation %d is not on a tile bit", k);
                return LQ_ERR_DEVICE;
            }
        }
        meta[o] = m;
    }
    // runs of the backward passes (see AdjRun)
    std::vector<AdjRun> runs;
    std::vector<std::pair<int, int>> pass_runs;  // per backward step: first run, run count
    for (const Step& st : plb.steps) {
        const int r0 = (int)runs.size();
        if (!st.normalize) {
            AdjRun cur{0, 0, -1, 0};
            int cur_params = 0;
            bool open = false;
            for (int o = 0; o < st.pass.op_count; ++o) {
                const TileOp& op = plb.ops[st.pass.op_first + o];
                const int par = meta[st.pass.op_first + o].slot >= 0 ? 1 : 0;
                if (op.kind == kTSwap2 || op.kind == kTQuad) {
                    if (open) {
                        if (cur.lb < 0) cur.lb = 0;
                        runs.push_back(cur);
                        open = false;
                    }
                    runs.push_back(AdjRun{o, 1, -1, 0});
                    continue;
                }
                const bool pair = op.kind != kTDiag;
                if (open && (!pair || cur.lb < 0 || cur.lb == op.lb0) && cur_params + par <= kRunParams) {
                    cur.count++;
                    cur_params += par;
                    if (pair) cur.lb = op.lb0;
                    continue;
                }
                if (open) {
                    if (cur.lb < 0) cur.lb = 0;  // diagonal ops only: any pairing bit
                    runs.push_back(cur);
                }
                cur = AdjRun{o, 1, pair ? op.lb0 : -1, 0};
                cur_params = par;
                open = true;
            }
            if (open) {
                if (cur.lb < 0) cur.lb = 0;
                runs.push_back(cur);
            }
        }
        pass_runs.emplace_back(r0, (int)runs.size() - r0);
    }
    for (const Step& st : plf.steps)
        if (st.normalize) {
            set_error("adjoint gradient: renormalising step in the forward plan");
            return LQ_ERR_UNSUPPORTED;
        }
    jit_prepare(plf, dtype, (uint64_t)count);
    const uint64_t dim = 1ull << n;
    const size_t amp = sizeof(V);
    const uint64_t nwp = (dim >> Kb) * kAWaves;
    const int nbE = (int)std::min<uint64_t>(std::max<uint64_t>(dim / (kAThreads * 16), 1), 64);
    const size_t nof = plf.ops.size(), nob = plb.ops.size();
    // instances per chunk: device states and partials within half the free
    // memory (at most 64 GiB), host op tables within 512 MiB
    static size_t budget = 0;  // queried once: the device query is not free
    if (!budget) {
        size_t freeb = 0, totb = 0;
        if (hipMemGetInfo(&freeb, &totb) != hipSuccess) freeb = (size_t)8 << 30;
        budget = std::min<size_t>(freeb / 2, (size_t)64 << 30);
    }
    const size_t per_dev = 2 * dim * amp + (size_t)std::max(nslots, 1) * nwp * 8 + 4096;
    const size_t per_host = (nof + nob) * sizeof(TileOp) + (size_t)std::max(nt, 1) * sizeof(PauliTermDev) + 64;
    int chunk = (int)std::min<size_t>((size_t)count, std::max<size_t>(1, budget / per_dev));
    chunk = (int)std::min<size_t>((size_t)chunk, std::max<size_t>(1, ((size_t)512 << 20) / per_host));
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    std::vector<lq_gate> gi, ri(num_gates);
    std::vector<Prim> pfi, pbi;
    PauliSum psi_sum;
    for (int c0 = 0; c0 < count; c0 += chunk) {
        const int cn = std::min(chunk, count - c0);
        const size_t b_of = al(cn * nof * sizeof(TileOp) + 1), b_ob = al(cn * nob * sizeof(TileOp) + 1);
        const size_t b_gr = al(plf.groups.size() * sizeof(TileGroup) + 1);
        const size_t b_me = al(meta.size() * sizeof(AdjMeta) + 1) + al(runs.size() * sizeof(AdjRun) + 1);
        const size_t b_te = al((size_t)cn * std::max(nt, 1) * sizeof(PauliTermDev));
        const size_t b_pa = al((size_t)cn * std::max(nslots, 1) * nwp * sizeof(double));
        const size_t b_pe = al((size_t)cn * nbE * sizeof(double));
        const size_t b_re = al((size_t)cn * (std::max(nslots, 1) + 1) * sizeof(double));
        const size_t up = b_of + b_ob + b_gr + b_me + b_te;
        char* base = adj_buffer(up + b_pa + b_pe + b_re);
        unsigned char* hb = adj_host(up + b_re);
        void* sts = adj_states((size_t)2 * cn * dim * amp);
        if (!base || !hb || !sts) {
            set_error("adjoint buffer allocation failed");
            return LQ_ERR_OUT_OF_MEMORY;
        }
        // the staging buffer may still feed the previous chunk's copy
        LQ_HIP_CHECK(hipStreamSynchronize(stream()));
        TileOp* h_of = (TileOp*)hb;
        TileOp* h_ob = (TileOp*)(hb + b_of);
        PauliTermDev* h_te = (PauliTermDev*)(hb + b_of + b_ob + b_gr + b_me);
        for (int i = 0; i < cn; ++i) {
            const int inst = c0 + i;
            bind_template(tmpl, num_gates, params + (size_t)inst * num_params, num_params, gi);
            for (int k = 0; k < num_gates; ++k) inverse_gate(gi[k], ri[num_gates - 1 - k]);
            pfi.clear();  // gates_to_prims appends
            pbi.clear();
            rc = gates_to_prims(n, gi.data(), num_gates, pfi);
            if (!rc) rc = gates_to_prims(n, ri.data(), num_gates, pbi);
            if (!rc) rc = compile_pauli_sum(n, paulis, coeffs + (size_t)inst * num_terms, num_terms, convention,
                                            &psi_sum);
            if (rc) return rc;
            if (pfi.size() != pf.size() || pbi.size() != pb.size() || (int)psi_sum.terms.size() != nt) {
                set_error("adjoint batch: instance %d differs in structure", inst);
                return LQ_ERR_INVALID_ARGUMENT;
            }
            for (size_t o = 0; o < nof; ++o) {
                TileOp t = plf.ops[o];
                std::memcpy(t.m, pfi[fpf[plf.op_src[o]]].m, sizeof(t.m));
                h_of[(size_t)i * nof + o] = t;
            }
            for (size_t o = 0; o < nob; ++o) {
                TileOp t = plb.ops[o];
                std::memcpy(t.m, pbi[fpb[plb.op_src[o]]].m, sizeof(t.m));
                h_ob[(size_t)i * nob + o] = t;
            }
            if (nt) std::memcpy(h_te + (size_t)i * nt, psi_sum.terms.data(), nt * sizeof(PauliTermDev));
        }
        if (!plf.groups.empty())
            std::memcpy(hb + b_of + b_ob, plf.groups.data(), plf.groups.size() * sizeof(TileGroup));
        if (!meta.empty()) std::memcpy(hb + b_of + b_ob + b_gr, meta.data(), meta.size() * sizeof(AdjMeta));
        const size_t runs_off = b_of + b_ob + b_gr + al(meta.size() * sizeof(AdjMeta) + 1);
        if (!runs.empty()) std::memcpy(hb + runs_off, runs.data(), runs.size() * sizeof(AdjRun));
        LQ_HIP_CHECK(hipMemcpyAsync(base, hb, up, hipMemcpyHostToDevice, stream()));
        const TileOp* d_of = (const TileOp*)base;
        const TileOp* d_ob = (const TileOp*)(base + b_of);
        const TileGroup* d_gr = (const TileGroup*)(base + b_of + b_ob);
        const AdjMeta* d_me = (const AdjMeta*)(base + b_of + b_ob + b_gr);
        const AdjRun* d_ru = (const AdjRun*)(base + runs_off);
        const PauliTermDev* d_te = (const PauliTermDev*)(base + b_of + b_ob + b_gr + b_me);
        double* d_pa = (double*)(base + up);
        double* d_pe = (double*)(base + up + b_pa);
        double* d_res = (double*)(base + up + b_pa + b_pe);  // [cn * nslots] gradients, then [cn] energies
        V* psi = (V*)sts;
        V* lam = psi + (size_t)cn * dim;
        const uint64_t total = (uint64_t)cn * dim;
        // forward sweep from |0...0>
        hipLaunchKernelGGL(k_init_states<V>, dim3(grid_of(total)), dim3(kAThreads), 0, stream(), psi, dim, total, 1);
        LQ_LAUNCH_CHECK();
        for (const Step& st : plf.steps) {
            rc = launch_tile_pass(psi, dtype, st.pass, d_of, d_gr, cn, dim, nof, st.jit);
            if (rc) return rc;
        }
        // lambda = H psi, E = Re <psi|lambda>
        hipLaunchKernelGGL(k_init_states<V>, dim3(grid_of(total)), dim3(kAThreads), 0, stream(), lam, dim, total, 0);
        for (const PauliGroup& grp : ps0.groups)
            hipLaunchKernelGGL(k_pauli_apply_add<V>, dim3(grid_of(total)), dim3(kAThreads), 0, stream(),
                               (const V*)psi, lam, n, total, grp.x, d_te + grp.first, nt, grp.count);
        LQ_LAUNCH_CHECK();
        double* d_e = d_res + (size_t)cn * std::max(nslots, 1);
        if (energies) {
            hipLaunchKernelGGL(k_re_inner<V>, dim3(nbE, cn), dim3(kAThreads), 0, stream(), (const V*)psi,
                               (const V*)lam, dim, d_pe);
            hipLaunchKernelGGL(k_rows_sum, dim3(cn), dim3(kAThreads), 0, stream(), (const double*)d_pe,
                               (uint64_t)nbE, d_e);
            LQ_LAUNCH_CHECK();
        }
        // backward sweep
        for (size_t si = 0; si < plb.steps.size(); ++si) {
            const Step& st = plb.steps[si];
            if (st.normalize) continue;
            rc = dispatch_adj<V, T>(Kb, psi, lam, st.pass, d_ob, d_me, d_ru + pass_runs[si].first,
                                    pass_runs[si].second, cn, dim, nob, d_pa, nslots, nwp);
            if (rc) return rc;
        }
        if (nslots > 0) {
            hipLaunchKernelGGL(k_rows_sum, dim3(cn * nslots), dim3(kAThreads), 0, stream(), (const double*)d_pa, nwp,
                               d_res);
            LQ_LAUNCH_CHECK();
        }
        double* h_res = (double*)(hb + up);
        const size_t nres = (size_t)cn * (std::max(nslots, 1) + 1);
        LQ_HIP_CHECK(hipMemcpyAsync(h_res, d_res, nres * sizeof(double), hipMemcpyDeviceToHost, stream()));
        LQ_HIP_CHECK(hipStreamSynchronize(stream()));
        for (int i = 0; i < cn; ++i) {
            double* gr = grads + (size_t)(c0 + i) * num_params;
            for (int p = 0; p < num_params; ++p) gr[p] = 0.0;
            for (int k = 0; k < num_gates; ++k)
                if (slot_of[k] >= 0) gr[tmpl[k].param[0]] += h_res[(size_t)i * nslots + slot_of[k]];
            if (energies) energies[c0 + i] = h_res[(size_t)cn * std::max(nslots, 1) + i];
        }
    }
    return LQ_OK;
}
}  // namespace lq

using namespace lq;

static int adjoint_entry(int num_qubits, int dtype, const lq_gate* tmpl, int num_gates, int count,
                         const double* params, int num_params, const uint8_t* paulis, const double* coeffs,
                         int num_terms, int convention, double* energies, double* grads) {
    clear_error();
    if (num_qubits < 1 || num_qubits > kMaxQubits || num_params < 0 || count < 0 ||
        (count > 0 && num_params > 0 && (!params || !grads)) || (dtype != LQ_C128 && dtype != LQ_C64) ||
        num_terms < 0 || (num_terms > 0 && (!paulis || !coeffs))) {
        set_error("invalid arguments");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    if (count == 0) return LQ_OK;
    int rc = check_template(num_qubits, tmpl, num_gates, num_params);
    if (rc) return rc;
    rc = ensure_init();
    if (rc) return rc;
    if (dtype == LQ_C128)
        return adjoint_batch_t<double2, double>(num_qubits, dtype, tmpl, num_gates, count, params, num_params, paulis,
                                                coeffs, num_terms, convention, energies, grads);
    return adjoint_batch_t<float2, float>(num_qubits, dtype, tmpl, num_gates, count, params, num_params, paulis,
                                          coeffs, num_terms, convention, energies, grads);
}

extern "C" int lq_adjoint_gradient(int num_qubits, int dtype, const lq_gate* tmpl, int num_gates,
                                   const double* params, int num_params, const uint8_t* paulis, const double* coeffs,
                                   int num_terms, int convention, double* energy_out, double* grad_out) {
    LQ_API_GUARD;
    std::vector<double> dummy;
    if (num_params == 0) dummy.resize(1);
    return adjoint_entry(num_qubits, dtype, tmpl, num_gates, 1, params, num_params, paulis, coeffs, num_terms,
                         convention, energy_out, num_params ? grad_out : dummy.data());
}

extern "C" int lq_adjoint_gradient_batch(int num_qubits, int dtype, const lq_gate* tmpl, int num_gates, int count,
                                         const double* params, int num_params, const uint8_t* paulis,
                                         const double* coeffs, int num_terms, int convention, double* energies,
                                         double* grads) {
    LQ_API_GUARD;
    return adjoint_entry(num_qubits, dtype, tmpl, num_gates, count, params, num_params, paulis, coeffs, num_terms,
                         convention, energies, grads);
}
