// Fused tile-pass engine for Circuit::execute (src/circuits/mod.rs:352-375).
//
// A pass streams the state through LDS in tiles of 2^K amplitudes.  A tile
// holds every combination of its K tile bits for one fixed value of the other
// n-K bits: the low L bits (contiguous in memory, so loads and stores are
// coalesced runs of 2^L amplitudes) plus K-L "high" bits chosen per pass.
// Every gate whose non-diagonal target bits are tile bits is applied in LDS,
// so one HBM read and one HBM write serve a whole run of gates.  Diagonal
// gates (RZ, CZ, CP, phases; controls anywhere) need no tile bits: each
// amplitude knows its global index.
//
// The planner walks the gate list in order and moves a gate into the current
// pass when its target bits fit and it commutes with every gate deferred
// before it (disjoint non-diagonal supports), so the result equals in-order
// execution up to floating-point rounding.  MatrixGate's renormalisation
// (src/gates/mod.rs:91) is a barrier: the pass ends, State::normalize runs,
// planning continues.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "lq_fused.h"
#include "lq_tile_ops.h"

namespace lq {

// ---------------------------------------------------------------------------------
// gate -> primitive
// ---------------------------------------------------------------------------------
static inline uint64_t bmask(int n, int q) { return 1ull << (n - 1 - q); }
static inline int bpos(int n, int q) { return n - 1 - q; }

static void polar2(double r, double th, double* out) {
    out[0] = r * std::cos(th);
    out[1] = r * std::sin(th);
}

static Prim make_prim(int kind) {
    Prim p;
    std::memset(&p, 0, sizeof(p));
    p.kind = kind;
    p.b0 = p.b1 = -1;
    p.sel0 = p.sel1 = -1;
    p.mclass = kTPairMat;
    return p;
}

static Prim pair_prim(int kind, int b, uint64_t ctrl, const double* m8) {
    Prim p = make_prim(kind);
    p.b0 = b;
    p.ctrl = ctrl;
    if (m8) std::memcpy(p.m, m8, 8 * sizeof(double));
    p.nd = 1ull << b;
    p.ds = ctrl;
    return p;
}

static Prim diag_prim(uint64_t fixed, int sel0, int sel1, const double* d8) {
    Prim p = make_prim(kTDiag);
    p.ctrl = fixed;
    p.sel0 = sel0;
    p.sel1 = sel1;
    std::memcpy(p.m, d8, 8 * sizeof(double));
    p.nd = 0;
    p.ds = fixed | (sel0 >= 0 ? 1ull << sel0 : 0) | (sel1 >= 0 ? 1ull << sel1 : 0);
    return p;
}

int gates_to_prims(int n, const lq_gate* gates, int count, std::vector<Prim>& out) {
    for (int gi = 0; gi < count; ++gi) {
        const lq_gate* g = &gates[gi];
        const size_t first = out.size();
        int rc = validate_gate(n, g);
        if (rc) return rc;
        const int* q = g->qubits;
        const double* th = g->theta;
        double m[8];
        double d[8] = {1, 0, 1, 0, 1, 0, 1, 0};
        bool renorm = (g->flags & LQ_GATE_FLAG_RENORMALIZE) != 0;
        switch (g->kind) {
            case LQ_GATE_X: out.push_back(pair_prim(kTPairSwap, bpos(n, q[0]), 0, nullptr)); break;
            case LQ_GATE_Y: out.push_back(pair_prim(kTPairNegY, bpos(n, q[0]), 0, nullptr)); break;
            case LQ_GATE_Z:
                d[0] = -1.0;
                d[1] = 0.0;
                out.push_back(diag_prim(bmask(n, q[0]), -1, -1, d));
                out.back().dclass = kTDiagNeg;
                break;
            case LQ_GATE_H: {
                double s = 1.0 / std::sqrt(2.0);  // single_qubit.rs:237
                double hm[8] = {s, 0, s, 0, s, 0, -s, 0};
                out.push_back(pair_prim(kTPairH, bpos(n, q[0]), 0, hm));
                break;
            }
            case LQ_GATE_S:
                d[0] = 0.0;
                d[1] = 1.0;
                out.push_back(diag_prim(bmask(n, q[0]), -1, -1, d));
                break;
            case LQ_GATE_T:
                polar2(1.0, M_PI / 4.0, d);
                out.push_back(diag_prim(bmask(n, q[0]), -1, -1, d));
                break;
            case LQ_GATE_RX:
            case LQ_GATE_RY:
            case LQ_GATE_U3:
            case LQ_GATE_CRX:
            case LQ_GATE_CRY: {
                bool ctl = g->kind == LQ_GATE_CRX || g->kind == LQ_GATE_CRY;
                gate_pair_matrix(g->kind, th, m);
                Prim pp = pair_prim(kTPairMat, bpos(n, ctl ? q[1] : q[0]), ctl ? bmask(n, q[0]) : 0, m);
                // the structure follows from the kind, for every angle
                if (g->kind == LQ_GATE_RY || g->kind == LQ_GATE_CRY) pp.mclass = kTPairReal;
                if (g->kind == LQ_GATE_RX || g->kind == LQ_GATE_CRX) pp.mclass = kTPairRX;
                out.push_back(pp);
                break;
            }
            case LQ_GATE_RZ:
                polar2(1.0, -th[0] / 2.0, d);
                polar2(1.0, th[0] / 2.0, d + 2);
                out.push_back(diag_prim(0, bpos(n, q[0]), -1, d));
                break;
            case LQ_GATE_PHASE:
                polar2(1.0, th[0], d);
                out.push_back(diag_prim(bmask(n, q[0]), -1, -1, d));
                break;
            case LQ_GATE_CNOT: out.push_back(pair_prim(kTPairSwap, bpos(n, q[1]), bmask(n, q[0]), nullptr)); break;
            case LQ_GATE_SWAP: {
                Prim p = make_prim(kTSwap2);
                p.b0 = bpos(n, q[0]);
                p.b1 = bpos(n, q[1]);
                p.nd = (1ull << p.b0) | (1ull << p.b1);
                out.push_back(p);
                break;
            }
            case LQ_GATE_CZ:
                d[0] = -1.0;
                d[1] = 0.0;
                out.push_back(diag_prim(bmask(n, q[0]) | bmask(n, q[1]), -1, -1, d));
                out.back().dclass = kTDiagNeg;
                break;
            case LQ_GATE_CPHASE:
            case LQ_GATE_CP:
                polar2(1.0, th[0], d);
                out.push_back(diag_prim(bmask(n, q[0]) | bmask(n, q[1]), -1, -1, d));
                break;
            case LQ_GATE_CRZ:
                polar2(1.0, -th[0] / 2.0, d);
                polar2(1.0, th[0] / 2.0, d + 2);
                out.push_back(diag_prim(bmask(n, q[0]), bpos(n, q[1]), -1, d));
                break;
            case LQ_GATE_RXX:
            case LQ_GATE_RYY: {
                Prim p = make_prim(kTQuad);
                p.b0 = bpos(n, q[0]);
                p.b1 = bpos(n, q[1]);
                two_qubit_rot_matrix(g->kind, th[0], q[0], q[1], p.m);
                p.nd = (1ull << p.b0) | (1ull << p.b1);
                out.push_back(p);
                break;
            }
            case LQ_GATE_RZZ:
                polar2(1.0, -th[0] / 2.0, d);
                polar2(1.0, th[0] / 2.0, d + 2);
                polar2(1.0, th[0] / 2.0, d + 4);
                polar2(1.0, -th[0] / 2.0, d + 6);
                out.push_back(diag_prim(0, bpos(n, q[0]), bpos(n, q[1]), d));
                break;
            case LQ_GATE_TOFFOLI:
                out.push_back(pair_prim(kTPairSwap, bpos(n, q[2]), bmask(n, q[0]) | bmask(n, q[1]), nullptr));
                break;
            case LQ_GATE_MATRIX1:
                out.push_back(pair_prim(kTPairMat, bpos(n, q[0]), 0, g->matrix));
                renorm = true;
                break;
            case LQ_GATE_MATRIX2: {
                int q1 = std::min(q[0], q[1]), q2 = std::max(q[0], q[1]);
                Prim p = make_prim(kTQuad);
                p.b0 = bpos(n, q1);
                p.b1 = bpos(n, q2);
                std::memcpy(p.m, g->matrix, 32 * sizeof(double));
                p.nd = (1ull << p.b0) | (1ull << p.b1);
                out.push_back(p);
                renorm = true;
                break;
            }
            default:
                set_error("unknown gate kind %d", g->kind);
                return LQ_ERR_INVALID_ARGUMENT;
        }
        if (renorm) out.push_back(make_prim(-1));
        for (size_t k = first; k < out.size(); ++k) out[k].src = gi;
    }
    return LQ_OK;
}

// ---------------------------------------------------------------------------------
// planner
// ---------------------------------------------------------------------------------
// Tile geometry: 2^12 complex128 / 2^13 complex64 amplitudes (64 KiB of LDS,
// two tiles per CU) with 2^5-amplitude contiguous runs unless overridden.
static int g_tile_c128 = 12, g_tile_c64 = 13, g_low = 5;
static bool g_c64_explicit = false;  // the complex64 tile size was set by the caller

int default_tile_bits(int n, int dtype) {
    int k = dtype == LQ_C128 ? g_tile_c128 : g_tile_c64;
    return std::min(k, n);
}

// Batches of complex64 states take 2^12-amplitude tiles unless the caller set
// a size: 32 KiB per block leaves room for four blocks per CU and blocks of
// 256 threads (which read the first group one tile ahead); their passes
// carry many gates (the ansatz layers), which that occupancy serves better
// than fewer, larger passes (512-circuit uneven batch 17.9 -> 14.8 ms,
// 20-qubit batched gradient 24.0 -> 21.3 ms), while one large state with few
// gates per pass keeps 2^13 (30-qubit brickwork 126 against 133 ms).
int batch_tile_bits(int n, int dtype) {
    if (dtype == LQ_C64 && !g_c64_explicit) return std::min(12, n);
    return default_tile_bits(n, dtype);
}

int default_low_bits(int n, int dtype) {
    (void)dtype;
    return std::min(g_low, n);
}

static TileOp to_tile_op(const Prim& p, int L, const int* hpos, int hb) {
    auto local = [&](int b) -> int32_t {
        if (b < 0) return -1;
        if (b < L) return b;
        for (int j = 0; j < hb; ++j)
            if (hpos[j] == b) return L + j;
        return -1;
    };
    TileOp t;
    std::memset(&t, 0, sizeof(t));
    t.kind = p.kind;
    t.lb0 = local(p.b0);
    t.lb1 = local(p.b1);
    t.sel0 = p.sel0;
    t.sel1 = p.sel1;
    t.ctrl = p.ctrl;
    t.mclass = p.mclass;
    t.dclass = p.dclass;
    std::memcpy(t.m, p.m, sizeof(t.m));
    return t;
}

// Partition a pass's ops into groups: maximal runs whose non-diagonal local
// target bits fit kGroupBits bits (diagonal ops join any group).  Free
// subcube positions go to the bits the group's ops test (controls,
// selectors), so those tests become compile-time per subcube slot; then to
// the highest free tile bits (lanes walk the low bits).  Each op gets its
// in-subcube / out-of-subcube control and selector split, and 2x2 matrices
// are classified (real, RX-like, general) for the leanest arithmetic.
static void build_groups(Plan& plan, TilePass& pass, int K) {
    pass.group_first = (int)plan.groups.size();
    pass.group_count = 0;
    pass.rich = 0;
    if (K < 8) return;  // small tiles use the op-by-op kernel
    auto local_of = [&](int g) -> int {
        if (g < pass.L) return g;
        for (int j = 0; j < pass.hb; ++j)
            if (pass.hpos[j] == g) return pass.L + j;
        return -1;
    };
    auto need = [&](const TileOp& t) -> uint32_t {
        if (t.kind == kTDiag) return 0u;
        uint32_t m = 1u << t.lb0;
        if (t.kind == kTSwap2 || t.kind == kTQuad) m |= 1u << t.lb1;
        return m;
    };
    int i = 0;
    while (i < pass.op_count) {
        uint32_t S = 0;
        int j = i;
        while (j < pass.op_count) {
            uint32_t nb = need(plan.ops[pass.op_first + j]);
            if (__builtin_popcount(S | nb) > kGroupBits) break;
            S |= nb;
            ++j;
        }
        // local bits tested by the group's ops, by frequency
        int freq[32] = {0};
        for (int o = i; o < j; ++o) {
            const TileOp& t = plan.ops[pass.op_first + o];
            for (int b = 0; b < 64; ++b)
                if (t.ctrl & (1ull << b)) {
                    int lb = local_of(b);
                    if (lb >= 0) freq[lb]++;
                }
            for (int sb : {t.sel0, t.sel1})
                if (sb >= 0) {
                    int lb = local_of(sb);
                    if (lb >= 0) freq[lb]++;
                }
        }
        while (__builtin_popcount(S) < kGroupBits) {
            int best = -1;
            for (int b = K - 1; b >= 0; --b)
                if (!(S & (1u << b)) && freq[b] > 0 && (best < 0 || freq[b] > freq[best])) best = b;
            if (best < 0) break;
            S |= 1u << best;
        }
        for (int b = K - 1; b >= 0 && __builtin_popcount(S) < kGroupBits; --b) S |= 1u << b;
        TileGroup g;
        std::memset(&g, 0, sizeof(g));
        g.op_first = i;
        g.op_count = j - i;
        int k = 0;
        uint64_t Sg = 0;
        int spos_of[32];
        for (int b = 0; b < 32; ++b) spos_of[b] = -1;
        for (int b = 0; b < K; ++b)
            if (S & (1u << b)) {
                g.spos[k] = b;
                g.gbit[k] = 1ull << (b < pass.L ? b : pass.hpos[b - pass.L]);
                Sg |= g.gbit[k];
                spos_of[b] = k;
                ++k;
            }
        for (int o = i; o < j; ++o) {
            TileOp& t = plan.ops[pass.op_first + o];
            t.sp0 = t.lb0 >= 0 ? spos_of[t.lb0] : -1;
            t.sp1 = t.lb1 >= 0 ? spos_of[t.lb1] : -1;
            t.cout = t.ctrl & ~Sg;
            t.cin = 0;
            for (int q = 0; q < kGroupBits; ++q)
                if (t.ctrl & g.gbit[q]) t.cin |= 1 << q;
            t.s0p = t.s1p = -1;
            t.osel0 = t.osel1 = -1;
            if (t.sel0 >= 0) {
                int lb = local_of(t.sel0);
                if (lb >= 0 && spos_of[lb] >= 0) t.s0p = spos_of[lb];
                else t.osel0 = t.sel0;
            }
            if (t.sel1 >= 0) {
                int lb = local_of(t.sel1);
                if (lb >= 0 && spos_of[lb] >= 0) t.s1p = spos_of[lb];
                else t.osel1 = t.sel1;
            }
            if (t.kind == kTPairMat) t.kind = t.mclass;
        }
        for (int o = i; o < j; ++o) {
            const int kd = plan.ops[pass.op_first + o].kind;
            if (kd == kTSwap2 || kd == kTQuad) pass.rich = 1;
        }
        plan.groups.push_back(g);
        pass.group_count++;
        i = j;
    }
}

// Order of a pass's ops for grouping: a list schedule that keeps every pair
// of non-commuting ops in their order (two ops commute when each acts only
// diagonally on the bits the other touches) and fills each register group
// before opening the next: all ready ops whose non-diagonal bits lie in the
// group's bits S are taken, then S grows by the ready op that adds the
// fewest bits (the earliest on ties), up to kGroupBits bits.  Returns
// positions into `take`.
static const bool g_group_order = !(std::getenv("LQ_GROUP_ORDER") && std::getenv("LQ_GROUP_ORDER")[0] == '0');

static std::vector<int> group_order(const std::vector<Prim>& prims, const std::vector<int>& take) {
    const int m = (int)take.size();
    std::vector<int> order;
    order.reserve(m);
    if (!g_group_order) {
        for (int i = 0; i < m; ++i) order.push_back(i);
        return order;
    }
    std::vector<std::vector<int>> succ(m);
    std::vector<int> npred(m, 0);
    for (int i = 0; i < m; ++i) {
        const Prim& a = prims[take[i]];
        for (int j = i + 1; j < m; ++j) {
            const Prim& b = prims[take[j]];
            if ((a.nd & (b.nd | b.ds)) || (b.nd & (a.nd | a.ds))) {
                succ[i].push_back(j);
                ++npred[j];
            }
        }
    }
    std::vector<char> done(m, 0);
    auto put = [&](int i) {
        done[i] = 1;
        order.push_back(i);
        for (int k : succ[i]) --npred[k];
    };
    while ((int)order.size() < m) {
        uint64_t S = 0;
        for (;;) {
            // successors have larger indices: one sweep takes every chain that fits
            for (int i = 0; i < m; ++i)
                if (!done[i] && npred[i] == 0 && (prims[take[i]].nd & ~S) == 0) put(i);
            int best = -1, best_add = 0;
            for (int i = 0; i < m; ++i) {
                if (done[i] || npred[i] != 0) continue;
                const uint64_t nd = prims[take[i]].nd;
                if (__builtin_popcountll(S | nd) > kGroupBits) continue;
                const int add = __builtin_popcountll(nd & ~S);
                if (best < 0 || add < best_add) {
                    best = i;
                    best_add = add;
                }
            }
            if (best < 0) break;
            S |= prims[take[best]].nd;
        }
    }
    return order;
}

// Pass high bits chosen among windows of consecutive bits as well as first
// fit (LQ_PLAN_SEARCH=0: first fit only).
static const bool g_plan_search = !(std::getenv("LQ_PLAN_SEARCH") && std::getenv("LQ_PLAN_SEARCH")[0] == '0');

void plan_passes(int n, const std::vector<Prim>& prims, int K, int L, Plan& plan) {
    plan.n = n;
    plan.steps.clear();
    plan.ops.clear();
    plan.op_src.clear();
    plan.groups.clear();
    plan.passes = 0;
    K = std::min(K, n);
    if (n <= K) {
        L = n;  // one tile is the whole state: fully contiguous
    }
    L = std::min(L, K);
    const int HB = K - L;
    const uint64_t lowmask = (L >= 64) ? ~0ull : ((1ull << L) - 1ull);
    size_t i = 0;
    while (i < prims.size()) {
        // segment up to the next renormalisation barrier
        size_t j = i;
        while (j < prims.size() && prims[j].kind >= 0) ++j;
        std::vector<int> pending;
        for (size_t k = i; k < j; ++k) pending.push_back((int)k);
        while (!pending.empty()) {
            uint64_t H = 0;
            uint64_t dnd = 0, dds = 0;
            std::vector<int> take, keep;
            for (int idx : pending) {
                const Prim& p = prims[idx];
                uint64_t need = p.nd & ~lowmask;
                bool blocked = (p.nd & (dnd | dds)) || (p.ds & dnd);
                if (!blocked && __builtin_popcountll(H | need) <= HB) {
                    H |= need;
                    take.push_back(idx);
                } else {
                    dnd |= p.nd;
                    dds |= p.ds;
                    keep.push_back(idx);
                }
            }
            // fill the high set: bits the deferred gates need first, then the
            // lowest free bits (longer contiguous runs)
            for (int idx : keep) {
                uint64_t need = prims[idx].nd & ~lowmask & ~H;
                if (need && __builtin_popcountll(H | need) <= HB) H |= need;
            }
            for (int b = L; b < n && __builtin_popcountll(H) < HB; ++b) H |= 1ull << b;
            if (g_plan_search && HB > 0 && n > K) {
                // a fixed high set takes every gate that fits it and is not
                // held back by a deferred one: keep the window of HB
                // consecutive high bits that takes the most, when it beats
                // the first-fit set (the first-fit set also takes gates
                // that its later-added bits would allow)
                auto take_with = [&](uint64_t Hs, std::vector<int>* t, std::vector<int>* k) {
                    uint64_t bn = 0, bs = 0;
                    size_t count = 0;
                    for (int idx : pending) {
                        const Prim& p = prims[idx];
                        const bool blocked = (p.nd & (bn | bs)) || (p.ds & bn);
                        if (!blocked && ((p.nd & ~lowmask) & ~Hs) == 0) {
                            ++count;
                            if (t) t->push_back(idx);
                        } else {
                            bn |= p.nd;
                            bs |= p.ds;
                            if (k) k->push_back(idx);
                        }
                    }
                    return count;
                };
                size_t best = take_with(H, nullptr, nullptr);
                uint64_t bestH = H;
                for (int b = L; b + HB <= n; ++b) {
                    const uint64_t Hw = ((HB >= 64 ? ~0ull : ((1ull << HB) - 1ull)) << b);
                    const size_t c = take_with(Hw, nullptr, nullptr);
                    if (c > best) {
                        best = c;
                        bestH = Hw;
                    }
                }
                if (bestH != H || take.size() != best) {
                    H = bestH;
                    take.clear();
                    keep.clear();
                    take_with(H, &take, &keep);
                }
            }
            Step st;
            st.normalize = false;
            std::memset(&st.pass, 0, sizeof(st.pass));
            st.pass.L = L;
            st.pass.hb = 0;
            for (int b = L; b < n; ++b)
                if (H & (1ull << b)) st.pass.hpos[st.pass.hb++] = b;
            st.pass.op_first = (int)plan.ops.size();
            st.pass.op_count = (int)take.size();
            st.pass.ntiles = 1ull << (n - L - st.pass.hb);
            for (int pos : group_order(prims, take)) {
                const int idx = take[pos];
                plan.ops.push_back(to_tile_op(prims[idx], L, st.pass.hpos, st.pass.hb));
                plan.op_src.push_back(prims[idx].src);
            }
            build_groups(plan, st.pass, K);
            plan.steps.push_back(st);
            plan.passes++;
            pending.swap(keep);
        }
        if (j < prims.size()) {
            Step st;
            std::memset(&st, 0, sizeof(st));
            st.normalize = true;
            plan.steps.push_back(st);
        }
        i = j + 1;
    }
}

// ---------------------------------------------------------------------------------
// kernel
// ---------------------------------------------------------------------------------
constexpr int kTileThreads = 256;


template <typename V, typename T>
__device__ __forceinline__ V cm(const double* m, int i, V a) {
    return cmul((T)m[2 * i], (T)m[2 * i + 1], a);
}

// Select one of four uniform phases by two per-lane bits without indexing
// memory with a lane-dependent value.
template <typename V, typename T>
__device__ __forceinline__ V apply_phase(const double* m, int b0, int b1, V x) {
    const T r0 = (T)m[0], i0 = (T)m[1], r1 = (T)m[2], i1 = (T)m[3];
    const T r2 = (T)m[4], i2 = (T)m[5], r3 = (T)m[6], i3 = (T)m[7];
    T pr = b1 ? (b0 ? r3 : r2) : (b0 ? r1 : r0);
    T pi = b1 ? (b0 ? i3 : i2) : (b0 ? i1 : i0);
    return cmul(pr, pi, x);
}

// Threads per tile block: at most 16 amplitudes per thread.
__host__ __device__ constexpr int tile_threads(int K) { return K <= 12 ? 256 : (1 << (K - 4)); }

template <int K, typename V, typename T>
__global__ void __launch_bounds__(tile_threads(K)) k_tile(V* __restrict__ psi, TilePass p,
                                                         const TileOp* __restrict__ ops, uint64_t ninst,
                                                         uint64_t inst_stride, uint64_t ops_stride) {
    constexpr uint32_t TS = 1u << K;
    constexpr int NT = tile_threads(K);
    constexpr bool FULL = TS >= (uint32_t)NT;  // every thread owns TS/NT elements
    constexpr int EPT = FULL ? (int)(TS / NT) : 1;
    extern __shared__ __align__(16) unsigned char smem[];
    V* tile = reinterpret_cast<V*>(smem);
    uint64_t* hoff = reinterpret_cast<uint64_t*>(smem + TS * sizeof(V));
    const int L = p.L;
    const uint32_t lowmask = (1u << L) - 1u;
    const uint32_t tid = threadIdx.x;
    const bool own = FULL || tid < TS;
    for (uint32_t h = tid; h < (1u << p.hb); h += NT) {
        uint64_t o = 0;
        for (int j = 0; j < p.hb; ++j)
            if ((h >> j) & 1u) o |= 1ull << p.hpos[j];
        hoff[h] = o;
    }
    __syncthreads();
    const uint64_t total = ninst * p.ntiles;
    for (uint64_t w = blockIdx.x; w < total; w += gridDim.x) {
        // wave-uniform by construction; readfirstlane lets the compiler see it,
        // so op parameters are fetched with scalar loads
        const uint32_t inst = __builtin_amdgcn_readfirstlane((uint32_t)(w / p.ntiles));
        const uint64_t t = w - (uint64_t)inst * p.ntiles;
        V* s = psi + (uint64_t)inst * inst_stride;
        const TileOp* op = ops + (uint64_t)inst * ops_stride + p.op_first;
        uint64_t base = t << L;
        for (int j = 0; j < p.hb; ++j) base = insert_zero(base, p.hpos[j]);
        // global index of element e of this tile
        auto gidx = [&](uint32_t e) -> uint64_t { return base | (e & lowmask) | hoff[e >> L]; };
        // load: every load is in flight before the first LDS write
        if (own) {
            V r[EPT];
#pragma unroll
            for (int j = 0; j < EPT; ++j) r[j] = s[gidx(tid + (uint32_t)j * NT)];
#pragma unroll
            for (int j = 0; j < EPT; ++j) tile[tid + (uint32_t)j * NT] = r[j];
        }
        __syncthreads();
        for (int o = 0; o < p.op_count;) {
            const int kind = op[o].kind;
            if (kind == kTDiag) {
                // a run of diagonal ops: elements stay in registers across the run
                int o2 = o + 1;
                while (o2 < p.op_count && op[o2].kind == kTDiag) ++o2;
                if (own) {
                    // real and imaginary parts in separate scalar arrays so the
                    // run stays in registers
                    T xr[EPT], xi[EPT];
#pragma unroll
                    for (int j = 0; j < EPT; ++j) {
                        V v = tile[tid + (uint32_t)j * NT];
                        xr[j] = v.x;
                        xi[j] = v.y;
                    }
                    for (int d = o; d < o2; ++d) {
                        const uint64_t ctrl = op[d].ctrl;
                        const int s0 = op[d].sel0, s1 = op[d].sel1;
                        const T r0 = (T)op[d].m[0], i0 = (T)op[d].m[1], r1 = (T)op[d].m[2], i1 = (T)op[d].m[3];
                        const T r2 = (T)op[d].m[4], i2 = (T)op[d].m[5], r3 = (T)op[d].m[6], i3 = (T)op[d].m[7];
#pragma unroll
                        for (int j = 0; j < EPT; ++j) {
                            const uint64_t g = gidx(tid + (uint32_t)j * NT);
                            const bool on = (g & ctrl) == ctrl;
                            const bool b0 = s0 >= 0 && ((g >> s0) & 1ull);
                            const bool b1 = s1 >= 0 && ((g >> s1) & 1ull);
                            const T pr = on ? (b1 ? (b0 ? r3 : r2) : (b0 ? r1 : r0)) : (T)1;
                            const T pi = on ? (b1 ? (b0 ? i3 : i2) : (b0 ? i1 : i0)) : (T)0;
                            const T yr = pr * xr[j] - pi * xi[j];
                            const T yi = pr * xi[j] + pi * xr[j];
                            xr[j] = on ? yr : xr[j];
                            xi[j] = on ? yi : xi[j];
                        }
                    }
#pragma unroll
                    for (int j = 0; j < EPT; ++j) {
                        V v;
                        v.x = xr[j];
                        v.y = xi[j];
                        tile[tid + (uint32_t)j * NT] = v;
                    }
                }
                o = o2;
            } else if (kind <= kTPairH) {
                const int lb = op[o].lb0;
                const uint64_t ctrl = op[o].ctrl;
                const T m0 = (T)op[o].m[0], m1 = (T)op[o].m[1], m2 = (T)op[o].m[2], m3 = (T)op[o].m[3];
                const T m4 = (T)op[o].m[4], m5 = (T)op[o].m[5], m6 = (T)op[o].m[6], m7 = (T)op[o].m[7];
                // TS/2 pairs: PPT per thread when TS/2 >= NT, else one pair
                // for the first TS/2 threads
                constexpr bool PFULL = TS / 2 >= (uint32_t)NT;
                constexpr int PPT = PFULL ? (int)(TS / 2 / NT) : 1;
#pragma unroll
                for (int j = 0; j < PPT; ++j) {
                    const uint32_t q = tid + (uint32_t)j * NT;
                    if (!PFULL && q >= TS / 2) continue;
                    const uint32_t e0 = ins0(q, lb);
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
                ++o;
            } else if (kind == kTSwap2) {
                const int lb0 = op[o].lb0, lb1 = op[o].lb1;
                const int lo = min(lb0, lb1), hi = max(lb0, lb1);
                for (uint32_t q = tid; q < TS / 4; q += NT) {
                    const uint32_t e = ins0(ins0(q, lo), hi) | (1u << lb0);
                    const uint32_t f = e ^ (1u << lb0) ^ (1u << lb1);
                    V x = tile[e];
                    tile[e] = tile[f];
                    tile[f] = x;
                }
                ++o;
            } else {  // kTQuad
                const TileOp& po = op[o];
                const int lo = min(po.lb0, po.lb1), hi = max(po.lb0, po.lb1);
                const uint32_t ma = 1u << po.lb0, mb = 1u << po.lb1;
                for (uint32_t q = tid; q < TS / 4; q += NT) {
                    const uint32_t e = ins0(ins0(q, lo), hi);
                    V a0 = tile[e], a1 = tile[e | mb], a2 = tile[e | ma], a3 = tile[e | ma | mb];
                    V y0 = cadd(cadd(cm<V, T>(po.m, 0, a0), cm<V, T>(po.m, 1, a1)),
                                cadd(cm<V, T>(po.m, 2, a2), cm<V, T>(po.m, 3, a3)));
                    V y1 = cadd(cadd(cm<V, T>(po.m, 4, a0), cm<V, T>(po.m, 5, a1)),
                                cadd(cm<V, T>(po.m, 6, a2), cm<V, T>(po.m, 7, a3)));
                    V y2 = cadd(cadd(cm<V, T>(po.m, 8, a0), cm<V, T>(po.m, 9, a1)),
                                cadd(cm<V, T>(po.m, 10, a2), cm<V, T>(po.m, 11, a3)));
                    V y3 = cadd(cadd(cm<V, T>(po.m, 12, a0), cm<V, T>(po.m, 13, a1)),
                                cadd(cm<V, T>(po.m, 14, a2), cm<V, T>(po.m, 15, a3)));
                    tile[e] = y0;
                    tile[e | mb] = y1;
                    tile[e | ma] = y2;
                    tile[e | ma | mb] = y3;
                }
                ++o;
            }
            __syncthreads();
        }
        if (own) {
            V r[EPT];
#pragma unroll
            for (int j = 0; j < EPT; ++j) r[j] = tile[tid + (uint32_t)j * NT];
#pragma unroll
            for (int j = 0; j < EPT; ++j) s[gidx(tid + (uint32_t)j * NT)] = r[j];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------
// grouped kernel (tiles of >= 2^8 amplitudes)
// ---------------------------------------------------------------------------------
// At least two waves per SIMD: two 64 KiB tiles per CU are resident, so the
// register budget is 256 VGPRs per lane.
template <int K, bool RICH, typename V, typename T>
__global__ void __launch_bounds__(tile_threads(K), 2) k_tileg(V* __restrict__ psi, TilePass p,
                                                          const TileOp* __restrict__ ops,
                                                          const TileGroup* __restrict__ groups, uint64_t ninst,
                                                          uint64_t inst_stride, uint64_t ops_stride) {
    static_assert(K >= 8, "grouped tiles need at least 256 amplitudes");
    constexpr uint32_t TS = 1u << K;
    constexpr int NT = tile_threads(K);
    constexpr int EPT = (int)(TS / NT);
    constexpr uint32_t NSUB = TS >> 4;
    constexpr bool SFULL = NSUB >= (uint32_t)NT;
    constexpr int SPT = SFULL ? (int)(NSUB / NT) : 1;
    extern __shared__ __align__(16) unsigned char smem[];
    V* tile = reinterpret_cast<V*>(smem);
    uint64_t* hoff = reinterpret_cast<uint64_t*>(smem + TS * sizeof(V));
    const int L = p.L;
    const uint32_t lowmask = (1u << L) - 1u;
    const uint32_t tid = threadIdx.x;
    for (uint32_t h = tid; h < (1u << p.hb); h += NT) {
        uint64_t o = 0;
        for (int j = 0; j < p.hb; ++j)
            if ((h >> j) & 1u) o |= 1ull << p.hpos[j];
        hoff[h] = o;
    }
    __syncthreads();
    const TileGroup* grp = groups + p.group_first;
    // a contiguous range of tiles per block: consecutive tiles share an
    // instance, so its op table stays in the scalar cache
    const uint64_t total = ninst * p.ntiles;
    const uint64_t per = (total + gridDim.x - 1) / gridDim.x;
    const uint64_t w0 = (uint64_t)blockIdx.x * per, w1 = w0 + per < total ? w0 + per : total;
    for (uint64_t w = w0; w < w1; ++w) {
        // wave-uniform by construction; readfirstlane lets the compiler see it,
        // so op parameters are fetched with scalar loads
        const uint32_t inst = __builtin_amdgcn_readfirstlane((uint32_t)(w / p.ntiles));
        const uint64_t t = w - (uint64_t)inst * p.ntiles;
        V* s = psi + (uint64_t)inst * inst_stride;
        const TileOp* op = ops + (uint64_t)inst * ops_stride + p.op_first;
        uint64_t base = t << L;
        for (int j = 0; j < p.hb; ++j) base = insert_zero(base, p.hpos[j]);
        auto gidx = [&](uint32_t e) -> uint64_t { return base | (e & lowmask) | hoff[e >> L]; };
        {
            V r[EPT];
#pragma unroll
            for (int j = 0; j < EPT; ++j) r[j] = s[gidx(tid + (uint32_t)j * NT)];
#pragma unroll
            for (int j = 0; j < EPT; ++j) tile[swz(tid + (uint32_t)j * NT)] = r[j];
        }
        __syncthreads();
        for (int gi = 0; gi < p.group_count; ++gi) {
            const TileGroup& G = grp[gi];
            const int s0 = G.spos[0], s1 = G.spos[1], s2 = G.spos[2], s3 = G.spos[3];
#pragma unroll
            for (int k = 0; k < SPT; ++k) {
                const uint32_t q = tid + (uint32_t)k * NT;
                if (!SFULL && q >= NSUB) break;
                const uint32_t eb = ins0(ins0(ins0(ins0(q, s0), s1), s2), s3);
                const uint64_t gb = gidx(eb);
                T xr[16], xi[16];
#pragma unroll
                for (int c = 0; c < 16; ++c) {
                    const uint32_t e = eb | ((uint32_t)(c & 1) << s0) | ((uint32_t)((c >> 1) & 1) << s1) |
                                       ((uint32_t)((c >> 2) & 1) << s2) | ((uint32_t)((c >> 3) & 1) << s3);
                    const V v = tile[swz(e)];
                    xr[c] = v.x;
                    xi[c] = v.y;
                }
                for (int o = G.op_first; o < G.op_first + G.op_count; ++o) g_apply<RICH, T>(xr, xi, op[o], gb);
#pragma unroll
                for (int c = 0; c < 16; ++c) {
                    const uint32_t e = eb | ((uint32_t)(c & 1) << s0) | ((uint32_t)((c >> 1) & 1) << s1) |
                                       ((uint32_t)((c >> 2) & 1) << s2) | ((uint32_t)((c >> 3) & 1) << s3);
                    V v;
                    v.x = xr[c];
                    v.y = xi[c];
                    tile[swz(e)] = v;
                }
            }
            __syncthreads();
        }
        {
            V r[EPT];
#pragma unroll
            for (int j = 0; j < EPT; ++j) r[j] = tile[swz(tid + (uint32_t)j * NT)];
#pragma unroll
            for (int j = 0; j < EPT; ++j) s[gidx(tid + (uint32_t)j * NT)] = r[j];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------
// launcher
// ---------------------------------------------------------------------------------
template <int K, bool RICH, typename V, typename T>
static void launch_g(void* psi, const TilePass& p, const TileOp* ops, const TileGroup* groups, uint64_t ninst,
                     uint64_t stride, uint64_t ostride, unsigned grid, size_t lds) {
    static std::once_flag once;
    std::call_once(once, [&] {
        (void)hipFuncSetAttribute((const void*)k_tileg<K, RICH, V, T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds + 4096);
    });
    hipLaunchKernelGGL((k_tileg<K, RICH, V, T>), dim3(grid), dim3(tile_threads(K)), lds, stream(), (V*)psi, p, ops,
                       groups, ninst, stride, ostride);
}

template <int K, typename V, typename T>
static void launch_k(void* psi, const TilePass& p, const TileOp* ops, const TileGroup* groups, uint64_t ninst,
                     uint64_t stride, uint64_t ostride, unsigned grid, size_t lds) {
    if constexpr (K >= 8) {
        if (p.rich) launch_g<K, true, V, T>(psi, p, ops, groups, ninst, stride, ostride, grid, lds);
        else launch_g<K, false, V, T>(psi, p, ops, groups, ninst, stride, ostride, grid, lds);
    } else {
        (void)groups;
        hipLaunchKernelGGL((k_tile<K, V, T>), dim3(grid), dim3(tile_threads(K)), lds, stream(), (V*)psi, p, ops,
                           ninst, stride, ostride);
    }
}

template <typename V, typename T>
static int dispatch_k(int K, void* psi, const TilePass& p, const TileOp* ops, const TileGroup* groups,
                      uint64_t ninst, uint64_t stride, uint64_t ostride, unsigned grid, size_t lds) {
    switch (K) {
#define LQ_K(k) \
    case k: launch_k<k, V, T>(psi, p, ops, groups, ninst, stride, ostride, grid, lds); break;
        LQ_K(1) LQ_K(2) LQ_K(3) LQ_K(4) LQ_K(5) LQ_K(6) LQ_K(7) LQ_K(8) LQ_K(9) LQ_K(10) LQ_K(11) LQ_K(12)
        LQ_K(13) LQ_K(14)
#undef LQ_K
        default:
            set_error("tile of 2^%d amplitudes is not supported", K);
            return LQ_ERR_UNSUPPORTED;
    }
    return LQ_OK;
}

int launch_tile_pass(void* psi, int dtype, const TilePass& p, const TileOp* d_ops, const TileGroup* d_groups,
                     uint64_t ninst, uint64_t inst_stride, uint64_t ops_stride, const JitKernel* jit) {
    const int K = p.L + p.hb;
    const size_t amp = dtype == LQ_C128 ? 16 : 8;
    const size_t lds = (amp << K) + (sizeof(uint64_t) << p.hb);
    uint64_t total = ninst * p.ntiles;
    unsigned grid = (unsigned)std::min<uint64_t>(total, 4096);
    if (jit) return launch_jit(jit, psi, p, d_ops, ninst, inst_stride, ops_stride, grid);
    int rc;
    if (dtype == LQ_C128)
        rc = dispatch_k<double2, double>(K, psi, p, d_ops, d_groups, ninst, inst_stride, ops_stride, grid, lds);
    else
        rc = dispatch_k<float2, float>(K, psi, p, d_ops, d_groups, ninst, inst_stride, ops_stride, grid, lds);
    if (rc) return rc;
    LQ_LAUNCH_CHECK();
    return LQ_OK;
}

// Device copy of a plan's ops and groups (grown on demand; reused across
// calls): [ops ...][groups ...].
static unsigned char* g_plan_dev = nullptr;
static unsigned char* g_plan_host = nullptr;
static size_t g_plan_cap = 0;

int upload_plan(const std::vector<TileOp>& ops, const std::vector<TileGroup>& groups, TileOp** d_ops,
                TileGroup** d_groups) {
    size_t ob = ops.size() * sizeof(TileOp);
    size_t gb = groups.size() * sizeof(TileGroup);
    size_t need = ob + gb + 256;
    LQ_HIP_CHECK(hipStreamSynchronize(stream()));  // the staging buffer may feed an earlier copy
    if (need > g_plan_cap) {
        if (g_plan_dev) (void)hipFree(g_plan_dev);
        if (g_plan_host) (void)hipHostFree(g_plan_host);
        g_plan_dev = nullptr;
        g_plan_host = nullptr;
        size_t cap = std::max<size_t>(need, 1 << 20);
        LQ_HIP_CHECK(hipMalloc((void**)&g_plan_dev, cap));
        LQ_HIP_CHECK(hipHostMalloc((void**)&g_plan_host, cap, hipHostMallocDefault));
        g_plan_cap = cap;
    }
    size_t goff = (ob + 255) & ~(size_t)255;
    if (ob) std::memcpy(g_plan_host, ops.data(), ob);
    if (gb) std::memcpy(g_plan_host + goff, groups.data(), gb);
    LQ_HIP_CHECK(hipMemcpyAsync(g_plan_dev, g_plan_host, goff + gb, hipMemcpyHostToDevice, stream()));
    *d_ops = (TileOp*)g_plan_dev;
    *d_groups = (TileGroup*)(g_plan_dev + goff);
    return LQ_OK;
}

int run_plan(lq_state* st, const Plan& plan, const TileOp* d_ops, const TileGroup* d_groups) {
    for (const Step& s : plan.steps) {
        int rc = s.normalize ? normalize_device(st, nullptr)
                             : launch_tile_pass(st->d_amps, st->dtype, s.pass, d_ops, d_groups, 1, 0, 0, s.jit);
        if (rc) return rc;
    }
    return LQ_OK;
}

int execute_fused(lq_state* st, const lq_gate* gates, int count) {
    std::vector<Prim> prims;
    int rc = gates_to_prims(st->n, gates, count, prims);
    if (rc) return rc;
    Plan plan;
    plan_passes(st->n, prims, default_tile_bits(st->n, st->dtype), default_low_bits(st->n, st->dtype), plan);
    if (plan.ops.empty() && plan.steps.empty()) return LQ_OK;
    jit_prepare(plan, st->dtype, 1);
    TileOp* d_ops = nullptr;
    TileGroup* d_groups = nullptr;
    rc = upload_plan(plan.ops, plan.groups, &d_ops, &d_groups);
    if (rc) return rc;
    return run_plan(st, plan, d_ops, d_groups);
}

}  // namespace lq

using namespace lq;

namespace lq {

int set_fused_config_internal(int tile_bits_c128, int tile_bits_c64, int low_bits) {
    int t1 = tile_bits_c128 > 0 ? tile_bits_c128 : 12;
    int t2 = tile_bits_c64 > 0 ? tile_bits_c64 : 13;
    int lo = low_bits > 0 ? low_bits : 5;
    // LDS: 16 B * 2^t1 and 8 B * 2^t2 must fit the 160 KiB per CU
    if (t1 > 13 || t2 > 14 || lo > t1 || lo > t2 || t1 < 1 || t2 < 1) {
        set_error("tile config out of range (c128 <= 13, c64 <= 14, low <= tile)");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    g_tile_c128 = t1;
    g_tile_c64 = t2;
    g_c64_explicit = tile_bits_c64 > 0;
    g_low = lo;
    return LQ_OK;
}

// The configuration as set_fused_config_internal takes it back (a complex64
// size the caller never set reads as 0, the default, so that a restore keeps
// batch_tile_bits's choice for batches).
void get_fused_config(int* tile_c128, int* tile_c64, int* low) {
    *tile_c128 = g_tile_c128;
    *tile_c64 = g_c64_explicit ? g_tile_c64 : 0;
    *low = g_low;
}

}  // namespace lq

extern "C" int lq_set_fused_config(int tile_bits_c128, int tile_bits_c64, int low_bits) {
    LQ_API_GUARD;
    clear_error();
    return set_fused_config_internal(tile_bits_c128, tile_bits_c64, low_bits);
}

extern "C" int lq_plan_fused(int num_qubits, int dtype, const lq_gate* gates, int count, int tile_bits,
                             int low_bits, int* gate_order, int* pass_index, int* num_passes) {
    LQ_API_GUARD;
    clear_error();
    if (num_qubits < 1 || num_qubits > kMaxQubits || count < 0 || (count > 0 && (!gates || !gate_order)) ||
        !num_passes) {
        set_error("invalid arguments");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    std::vector<Prim> prims;
    int rc = gates_to_prims(num_qubits, gates, count, prims);
    if (rc) return rc;
    Plan plan;
    int K = tile_bits > 0 ? tile_bits : default_tile_bits(num_qubits, dtype);
    int L = low_bits > 0 ? low_bits : default_low_bits(num_qubits, dtype);
    if (K > 14 || L > K) {
        set_error("tile_bits must be <= 14 and low_bits <= tile_bits");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    plan_passes(num_qubits, prims, K, L, plan);
    int k = 0, pass = 0;
    for (const Step& s : plan.steps) {
        if (s.normalize) continue;
        for (int o = 0; o < s.pass.op_count; ++o) {
            if (k >= count) break;
            gate_order[k] = plan.op_src[s.pass.op_first + o];
            if (pass_index) pass_index[k] = pass;
            ++k;
        }
        ++pass;
    }
    *num_passes = plan.passes;
    return LQ_OK;
}

namespace lq {

}  // namespace lq
